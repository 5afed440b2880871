// tgtp_batch.hip -- FAST_FP32 / FAST_BF16 scoring of one graph or of a batch
// of graphs, and the batched C ABI (tgtp_batch_*, tgtp_score_batch).
//
// A batch runs as the disjoint union of its graphs: one CSR (vertex ids
// shifted per graph), one eigenvector array, per-graph tables.  A FAST plan
// holds everything a run needs on the device -- per-graph scalars, the
// padded 16-edge tiles of the tau projection with their column tables, the
// blocks of 64 sources grouped into subsets whose n x 64 diffusion states
// fit the plan's state buffers (4 GiB each by default), the round bound of
// every subset, the hub rows of small graphs, and all scratch -- so a run is
// a fixed kernel sequence: clear, tau projection, per subset (init, rounds,
// transpose, post), symmetrization, with odd subsets on a second stream
// (its own buffer pair) when there are several.  A single graph is a batch
// of one, so the plain and batched forms compute the same scores bit for
// bit.  The batch captures its sequence into a hipGraph once and replays it;
// many small graphs then share each subset's launches instead of each
// paying its own.
//
// Autotuning (TGTP_BATCH_AUTOTUNE): coordinate descent over the degree up to
// which one thread scans a row in the post kernel, the bouquet count below
// which a walk position gets one 8-lane group instead of the whole
// workgroup, and the number of resident post workgroups, each candidate timed
// with HIP events on a full run; the best is kept for the captured graph.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "host_tgtp.hpp"
#include "taus_fast.hpp"
#include "fast_block.hpp"
#include "tgtp_fast.hpp"
#include "tgtp_internal.hpp"

namespace {

int env_int_batch(const char* name, int dflt) {
  const char* v = std::getenv(name);
  return v && *v ? std::atoi(v) : dflt;
}

struct GraphSpec {        // host description of one graph of a plan
  uint32_t voff, n;       // vertex range in the union
  uint64_t eoff, m;       // CSR range in the union
  const double* lam;      // [omega] host eigenvalues
  double eps, delta;
  int32_t gamma;
};

struct FastPlan {
  // identity (a plan cached on a single dgraph is reused when these match)
  uint64_t eig_gen = ~0ull;  // eigenpair upload of the graph the plan was built for
  double eps = 0, delta = 0;
  int32_t gamma = 0, mode = 0;
  uint32_t n_graphs = 0;
  uint64_t n = 0, m = 0;
  int32_t omega = 0;
  uint32_t omega_pad = 0;
  // device tables
  uint32_t* d_voff = nullptr;
  double* d_eps = nullptr;
  double* d_logmd = nullptr;
  int32_t* d_gamma = nullptr;
  TauGraph* d_tgraphs = nullptr;
  uint32_t* d_und = nullptr;
  uint32_t* d_tile_graph = nullptr;
  uint64_t n_tiles = 0;
  float* d_wt = nullptr;
  __bf16* d_wtb = nullptr;
  __bf16* d_xb = nullptr;
  // source blocks (64 sources of one graph, largest graphs first) and the
  // subsets whose n x 64 states share the two state buffers
  BlockInfo* d_blocks = nullptr;
  std::vector<BlockInfo> h_blocks;  // host copy (profiling)
  std::vector<uint64_t> g_m;        // CSR entries per graph
  uint32_t n_blocks = 0;
  uint64_t n_sources = 0;
  std::vector<BlockSubset> subsets;
  std::vector<uint32_t> round_tiles;  // per subset, rounds + 1 entries: tiles round r needs
  std::vector<uint32_t> round_hubs;   // ... and hub rows round r needs
  uint32_t hub_deg = 0;               // rows above this degree go to hub workgroups
  uint32_t* d_hubs = nullptr;         // hub rows (union ids) per graph
  uint32_t* d_hub_off = nullptr;      // [G + 1]
  // scratch
  int32_t* d_taus = nullptr;
  double* d_preds = nullptr;
  uint64_t* d_nr = nullptr;
  float* d_sa = nullptr;           // state buffers
  float* d_sb = nullptr;
  uint8_t* d_active = nullptr;
  int32_t* d_ell = nullptr;
  float* d_colf = nullptr;         // gmax, gmin per column
  uint32_t* d_colu = nullptr;      // qsize, st_mx, st_mn, st_q per column; last per block
  unsigned long long* d_dsum = nullptr;
  uint32_t* d_words = nullptr;
  float* d_out = nullptr;
  uint32_t max_slots = 0, slots = 0;  // resident workgroups of the post kernel
  uint32_t per_cu_staged = 1;  // ... per CU with the full LDS buffer (what slots is tuned for)
  uint32_t tile_mode = 0;
  uint32_t round_form = 0;  // blk_round_kernel (0) or the four-rows-per-wave form (1, 2)
  // tuned parameters
  uint32_t light = 32, small_bq = 8;
  // captured run
  // captured run: one graph, or four (pre, even subsets, odd subsets, post)
  // for a plan on two streams
  hipGraph_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
  hipGraphExec_t exec[4] = {nullptr, nullptr, nullptr, nullptr};
  // batches: odd subsets on a second stream with their own state buffers,
  // so one subset's post kernel (latency-bound walks) overlaps the next
  // subset's rounds (bandwidth-bound)
  bool dual = false;
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  float* d_sa2 = nullptr;
  float* d_sb2 = nullptr;
  uint64_t state_floats = 0;
};

void free_plan(FastPlan* P) {
  if (!P) return;
  for (int k = 0; k < 4; ++k) {
    if (P->exec[k]) (void)hipGraphExecDestroy(P->exec[k]);
    if (P->graph[k]) (void)hipGraphDestroy(P->graph[k]);
  }
  if (P->side) (void)hipStreamDestroy(P->side);
  if (P->fork) (void)hipEventDestroy(P->fork);
  if (P->join) (void)hipEventDestroy(P->join);
  if (P->d_sa2) (void)hipFree(P->d_sa2);
  if (P->d_sb2) (void)hipFree(P->d_sb2);
  void* ptrs[] = {P->d_voff, P->d_eps, P->d_logmd, P->d_gamma, P->d_tgraphs, P->d_und,
                  P->d_tile_graph, P->d_wt, P->d_wtb, P->d_xb, P->d_blocks, P->d_taus,
                  P->d_hubs, P->d_hub_off,
                  P->d_preds, P->d_nr, P->d_sa, P->d_sb, P->d_active, P->d_ell, P->d_colf,
                  P->d_colu, P->d_dsum, P->d_words, P->d_out};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete P;
}

template <typename T>
int dev_alloc(tgtp_ctx* ctx, T** p, uint64_t count, const char* what) {
  if (hipMalloc(reinterpret_cast<void**>(p), std::max<uint64_t>(count, 1) * sizeof(T)) !=
      hipSuccess) {
    *p = nullptr;
    return set_error(ctx, TGTP_E_NOMEM, std::string("device allocation: ") + what);
  }
  return TGTP_OK;
}

template <typename T>
int dev_upload(tgtp_ctx* ctx, T** p, const std::vector<T>& host, const char* what) {
  int rc = dev_alloc(ctx, p, host.size(), what);
  if (rc) return rc;
  if (!host.empty())
    TGTP_HIP_CHECK(ctx, hipMemcpyAsync(*p, host.data(), host.size() * sizeof(T),
                                       hipMemcpyHostToDevice, ctx->stream));
  return TGTP_OK;
}

TauFastArgs plan_tau_args(const FastPlan* P, tgtp_dgraph* dg);

// Builds the device tables and scratch of a FAST run over `dg` (a single
// graph or a union) whose eigenvectors are already in dg->X.  With `ranges`
// (a single graph only) the source blocks cover just those half-open source
// ranges; taus still cover every edge.
int build_plan(tgtp_ctx* ctx, tgtp_dgraph* dg, const std::vector<GraphSpec>& specs, int32_t omega,
               int32_t mode, FastPlan** out,
               const std::vector<std::pair<uint32_t, uint32_t>>* ranges = nullptr) {
  auto* prep = static_cast<TgtpPrep*>(dg->prep);
  auto* P = new (std::nothrow) FastPlan();
  if (!P) return set_error(ctx, TGTP_E_NOMEM, "host allocation");
  auto fail = [&](int rc) {
    free_plan(P);
    return rc;
  };
  hipStream_t s = ctx->stream;
  const uint32_t G = static_cast<uint32_t>(specs.size());
  P->n_graphs = G;
  P->n = dg->n;
  P->m = dg->m;
  P->mode = mode;
  P->omega = omega;
  P->omega_pad = static_cast<uint32_t>((omega + 31) / 32 * 32);
  if (G == 1) {
    P->eps = specs[0].eps;
    P->delta = specs[0].delta;
    P->gamma = specs[0].gamma;
    P->eig_gen = dg->eig_gen;
  }
  // per-graph scalars
  std::vector<uint32_t> voff(G + 1);
  std::vector<double> g_eps(G), g_logmd(G), tau0p(3 * G);
  std::vector<int32_t> g_gamma(G);
  for (uint32_t k = 0; k < G; ++k) {
    const GraphSpec& S = specs[k];
    voff[k] = S.voff;
    g_eps[k] = S.eps;
    g_logmd[k] = std::log(S.m / S.delta);  // log(g.m / config.delta), aesc.h:498,520
    g_gamma[k] = S.gamma;
    const double l1 = S.lam[1];
    tau0p[3 * k] = static_cast<double>(static_cast<float>(S.eps / 2.0));
    tau0p[3 * k + 1] = 1.0 - (l1 * l1);
    tau0p[3 * k + 2] = std::log(1.0 / std::fabs(l1));
  }
  voff[G] = static_cast<uint32_t>(dg->n);
  int rc;
  double* d_tau0p = nullptr;
  uint32_t* d_tau0max = nullptr;
  if ((rc = dev_upload(ctx, &P->d_voff, voff, "graph offsets")) ||
      (rc = dev_upload(ctx, &P->d_eps, g_eps, "graph eps")) ||
      (rc = dev_upload(ctx, &P->d_logmd, g_logmd, "graph log(m/delta)")) ||
      (rc = dev_upload(ctx, &P->d_gamma, g_gamma, "graph gamma")) ||
      (rc = dev_upload(ctx, &d_tau0p, tau0p, "tau0 params")) ||
      (rc = dev_upload(ctx, &d_tau0max, std::vector<uint32_t>(G, 0u), "tau0 max")))
    return fail(rc);
  rc = launch_tau0_per_graph(ctx, dg->idx, dg->adj, dg->m, static_cast<uint32_t>(dg->n), P->d_voff,
                             G, d_tau0p, d_tau0max);
  std::vector<uint32_t> tau0max(G);
  if (!rc &&
      hipMemcpyAsync(tau0max.data(), d_tau0max, G * 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
      hipStreamSynchronize(s) == hipSuccess) {
  } else if (!rc) {
    rc = set_error(ctx, TGTP_E_HIP, "tau0 readback");
  }
  (void)hipFree(d_tau0p);
  (void)hipFree(d_tau0max);
  if (rc) return fail(rc);
  // column tables: row c of graph k = column c of W_k
  std::vector<TauGraph> tg(G);
  uint32_t rows_total = 0;
  for (uint32_t k = 0; k < G; ++k) {
    const uint32_t ncols = 1 + std::min<uint32_t>(tau0max[k] / 2 + 3, 1u << 16);
    tg[k].n_col_tiles = (ncols + 15) / 16;
    tg[k].wt_row0 = rows_total;
    rows_total += tg[k].n_col_tiles * 16;
  }
  std::vector<float> wt(static_cast<size_t>(rows_total) * P->omega_pad, 0.0f);
  for (uint32_t k = 0; k < G; ++k) {
    const GraphSpec& S = specs[k];
    const uint32_t rows = tg[k].n_col_tiles * 16;
    const std::vector<double> pw = tgtp_host::power_table(S.lam, omega, rows);
    for (uint32_t c = 0; c < rows; ++c) {
      float* row = wt.data() + static_cast<size_t>(tg[k].wt_row0 + c) * P->omega_pad;
      for (int i = 1; i < omega - 1; ++i)
        row[i] = c == 0 ? static_cast<float>(1 + S.lam[i])
                        : static_cast<float>(pw[static_cast<size_t>(c - 1) * omega + i] /
                                             (1 - S.lam[i]));
    }
    const double lw = S.lam[omega - 1];
    tg[k].sq1m_1 = tau0p[3 * k + 1];
    tg[k].b_1 = tau0p[3 * k + 2];
    tg[k].sq1m_w = 1.0 - (lw * lw);
    tg[k].b_w = std::log(1.0 / std::fabs(lw));
    tg[k].m = static_cast<float>(S.m);
    tg[k].eps = static_cast<float>(S.eps / 2.0);
  }
  if ((rc = dev_upload(ctx, &P->d_wt, wt, "tau column tables")) ||
      (rc = dev_upload(ctx, &P->d_tgraphs, tg, "tau graph params")))
    return fail(rc);
  if (tgtp_mode_base(mode) == TGTP_MODE_FAST_BF16) {
    if ((rc = dev_alloc(ctx, &P->d_wtb, wt.size(), "bf16 tables")) ||
        (rc = launch_to_bf16(ctx, P->d_wt, wt.size(), P->d_wtb)) ||
        (rc = dev_alloc(ctx, &P->d_xb, dg->n * omega, "bf16 eigenvectors")) ||
        (rc = launch_to_bf16(ctx, dg->X, dg->n * omega, P->d_xb)))
      return fail(rc);
  }
  // 16-edge tiles per graph: positions p with adj[p] >= row(p), padded
  std::vector<uint32_t> und, tile_graph;
  und.reserve(dg->m / 2 + 16 * G);
  const uint32_t* idx = prep->h_idx.data();
  const uint32_t* adj = prep->h_adj.data();
  for (uint32_t k = 0; k < G; ++k) {
    const GraphSpec& S = specs[k];
    const size_t start = und.size();
    for (uint64_t u = S.voff; u < S.voff + S.n; ++u)
      for (uint64_t p = idx[u]; p < idx[u + 1]; ++p)
        if (adj[p] >= u) und.push_back(static_cast<uint32_t>(p));
    while ((und.size() - start) % 16) und.push_back(0xFFFFFFFFu);
    tile_graph.insert(tile_graph.end(), (und.size() - start) / 16, k);
  }
  P->n_tiles = tile_graph.size();
  if ((rc = dev_upload(ctx, &P->d_und, und, "tau tiles")) ||
      (rc = dev_upload(ctx, &P->d_tile_graph, tile_graph, "tile graphs")))
    return fail(rc);
  if ((rc = dev_alloc(ctx, &P->d_taus, dg->m, "taus")) ||
      (rc = dev_alloc(ctx, &P->d_preds, dg->m, "preds")) ||
      (rc = dev_alloc(ctx, &P->d_nr, dg->m, "walk counts")) ||
      (rc = dev_alloc(ctx, &P->d_out, dg->m, "scores")) ||
      (rc = dev_alloc(ctx, &P->d_words, 16, "counters")))
    return fail(rc);
  // taus once, to bound the rounds each block can take (a source stops by
  // the round equal to the largest tau of its row, when no walk is left)
  if ((rc = launch_fill_i32(ctx, P->d_taus, P->m, 0)) ||
      (rc = launch_taus_fast(ctx, plan_tau_args(P, dg),
                             tgtp_mode_base(mode) == TGTP_MODE_FAST_BF16)))
    return fail(rc);
  std::vector<int32_t> taus(dg->m);
  if (dg->m)
    TGTP_HIP_CHECK(ctx, hipMemcpyAsync(taus.data(), P->d_taus, dg->m * 4, hipMemcpyDeviceToHost, s));
  TGTP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  // blocks: graphs whose vectors the post kernel reads in place (n >
  // kPostLds), largest m first, then those it stages in LDS by decreasing
  // n, so that a subset's LDS buffer (sized for its largest staged vector)
  // fits its graphs closely and a subset seldom mixes the two kinds
  std::vector<uint32_t> order(G);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    const bool la = specs[a].n > kPostLds, lb = specs[b].n > kPostLds;
    if (la != lb) return la;
    return la ? specs[a].m > specs[b].m : specs[a].n > specs[b].n;
  });
  std::vector<BlockInfo> blocks;
  std::vector<uint32_t> bround;
  uint64_t n_sources = 0;
  for (uint32_t k : order) {
    const GraphSpec& S = specs[k];
    const std::vector<std::pair<uint32_t, uint32_t>> all{{0u, S.n}};
    for (const auto& r : ranges ? *ranges : all) {
      for (uint32_t s0 = r.first; s0 < r.second; s0 += kBlkCols) {
        BlockInfo B{};
        B.voff = S.voff;
        B.n = S.n;
        B.src0 = s0;
        B.cnt = std::min<uint32_t>(kBlkCols, r.second - s0);
        B.gi = k;
        int32_t tmax = 1;
        for (uint32_t v = S.voff + s0; v < S.voff + s0 + B.cnt; ++v)
          for (uint64_t p = idx[v]; p < idx[v + 1]; ++p) tmax = std::max(tmax, taus[p]);
        blocks.push_back(B);
        bround.push_back(static_cast<uint32_t>(tmax));
        n_sources += B.cnt;
      }
    }
  }
  P->n_blocks = static_cast<uint32_t>(blocks.size());
  P->n_sources = n_sources;
  // tile order: per-XCD runs of whole blocks, or (a block's state larger
  // than 16 MiB, eight of which would not fit the Infinity Cache) groups of
  // 8 tiles of one block across the XCDs
  uint32_t n_max = 0;
  for (const BlockInfo& B : blocks) n_max = std::max(n_max, B.n);
  P->tile_mode = static_cast<uint64_t>(n_max) * kBlkCols * 4 > (16ull << 20) ? 1u : 0u;
  const int tm = env_int_batch("TGTP_FAST_TILE_MODE", -1);
  if (tm == 0 || tm == 1) P->tile_mode = static_cast<uint32_t>(tm);
  // diffusion rounds four rows per wave: 16 gathers per step at 4 waves per
  // SIMD where rows are long (mean degree >= 8), 8 per step at 6 waves per
  // SIMD on short rows (grids); TGTP_FAST_ROUND=0 selects one row per wave
  P->round_form = dg->m >= 8ull * dg->n ? 2u : 1u;
  const int sb = env_int_batch("TGTP_FAST_SMALL_BQ", -1);
  if (sb >= 0) P->small_bq = static_cast<uint32_t>(sb);
  const int rf = env_int_batch("TGTP_FAST_ROUND", -1);
  if (rf >= 0 && rf <= 2) P->round_form = static_cast<uint32_t>(rf);
  // the four-rows form addresses a block's state and the adjacency with
  // 32-bit buffer offsets
  if (n_max >= (1u << 24) || dg->m >= (1ull << 30)) P->round_form = 0;
  // On small graphs (n <= 8192: at most (n/64)^2 = 16K tiles a round) rows
  // above 128 neighbours (8 steps of 16 in one quarter-wave) are summed by a
  // workgroup each in 16 segments: a hub's chain of dependent steps
  // otherwise sets the length of every round (n=4096: -17% per score).  On
  // larger graphs the other tiles hide the chains and the extra workgroups
  // cost more than they save (n=28,351, mean degree 34: +10%).
  // (in every round form: a graph's scores must not depend on which form
  // its plan picked, alone or in a batch; hub rows address a block's state
  // with 32-bit offsets, hence n < 2^24, and their own row with 32-bit byte
  // offsets, hence a graph with a row of 2^29 or more entries keeps all its
  // rows in the round kernels)
  P->hub_deg = 128u;
  const int hd = env_int_batch("TGTP_FAST_HUB_DEG", -1);
  if (hd >= 0) P->hub_deg = static_cast<uint32_t>(hd);
  const uint32_t hub_max_n = static_cast<uint32_t>(
      std::min(std::max(0, env_int_batch("TGTP_FAST_HUB_MAX_N", 8192)), (1 << 24) - 1));
  std::vector<uint32_t> hubs, hub_off(G + 1, 0);
  for (uint32_t k = 0; k < G; ++k) {
    hub_off[k] = static_cast<uint32_t>(hubs.size());
    if (!P->hub_deg || specs[k].n > hub_max_n) continue;
    bool short_rows = true;
    for (uint64_t v = specs[k].voff; v < specs[k].voff + specs[k].n; ++v) {
      const uint64_t d = idx[v + 1] - idx[v];
      short_rows = short_rows && d < (1ull << 29);
      if (d > P->hub_deg) hubs.push_back(static_cast<uint32_t>(v));
    }
    if (!short_rows) hubs.resize(hub_off[k]);
  }
  hub_off[G] = static_cast<uint32_t>(hubs.size());
  std::vector<uint32_t> hub_cnt(G);
  for (uint32_t k = 0; k < G; ++k) hub_cnt[k] = hub_off[k + 1] - hub_off[k];
  if ((rc = dev_upload(ctx, &P->d_hubs, hubs, "hub rows")) ||
      (rc = dev_upload(ctx, &P->d_hub_off, hub_off, "hub offsets")))
    return fail(rc);
  auto units_of = [&](uint32_t n) {
    const uint32_t t = (n + kBlkRows - 1) / kBlkRows;
    return P->tile_mode == 0 ? t : (t + 7) / 8;
  };
  // subsets of blocks whose states fit one state buffer (4 GiB by default:
  // measured against 1, 2, 8 and 16 GiB on the batched workload's large
  // graphs; larger subsets mean fewer post-kernel tails, beyond 4 GiB the
  // two streams' working sets crowd the caches)
  const uint64_t budget =
      static_cast<uint64_t>(std::max(1, env_int_batch("TGTP_FAST_STATE_MIB", 4096))) << 18;  // floats
  uint64_t max_floats = 0;
  for (uint32_t b0 = 0; b0 < P->n_blocks;) {
    BlockSubset S{};
    S.b0 = b0;
    uint64_t fl = 0;
    uint32_t tiles = 0, rounds = 1, hubs_before = 0;
    uint32_t b = b0;
    while (b < P->n_blocks) {
      const uint64_t need = static_cast<uint64_t>(blocks[b].n) * kBlkCols;
      if (b > b0 && fl + need > budget) break;
      fl += need;
      ++b;
    }
    // blocks that may need more rounds first: round r then covers a prefix
    // of the subset's tiles (the blocks whose round bound is >= r)
    std::vector<uint32_t> ord(b - b0);
    std::iota(ord.begin(), ord.end(), b0);
    std::stable_sort(ord.begin(), ord.end(),
                     [&](uint32_t x, uint32_t y) { return bround[x] > bround[y]; });
    std::vector<BlockInfo> bs;
    std::vector<uint32_t> br;
    for (uint32_t x : ord) {
      bs.push_back(blocks[x]);
      br.push_back(bround[x]);
    }
    fl = 0;
    for (uint32_t i = 0; i < bs.size(); ++i) {
      blocks[b0 + i] = bs[i];
      bround[b0 + i] = br[i];
      blocks[b0 + i].xoff = fl;
      blocks[b0 + i].tile0 = tiles;
      blocks[b0 + i].hub0 = hubs_before;
      blocks[b0 + i].hub_deg = hub_cnt[bs[i].gi] ? P->hub_deg : 0u;
      hubs_before += hub_cnt[bs[i].gi];
      fl += static_cast<uint64_t>(bs[i].n) * kBlkCols;
      tiles += units_of(bs[i].n);
      rounds = std::max(rounds, br[i]);
    }
    S.b1 = b;
    S.post_stage = 0;
    for (const BlockInfo& x : bs)
      if (x.n <= kPostLds) S.post_stage = std::max(S.post_stage, x.n);
    S.post_stage = std::min(kPostLds, (S.post_stage + 255u) / 256u * 256u);
    S.post_per_cu = static_cast<uint32_t>(block_post_per_cu(S.post_stage));
    S.tiles = tiles;
    S.rounds = rounds;
    S.floats = fl;
    S.round_tiles0 = static_cast<uint32_t>(P->round_tiles.size());
    for (uint32_t r = 0; r <= rounds; ++r) {
      uint32_t tr = 0, hr = 0;
      for (uint32_t i = 0; i < bs.size() && br[i] >= r; ++i) {
        tr = blocks[b0 + i].tile0 + units_of(bs[i].n);
        hr = blocks[b0 + i].hub0 + hub_cnt[bs[i].gi];
      }
      P->round_tiles.push_back(tr);
      P->round_hubs.push_back(hr);
    }
    max_floats = std::max(max_floats, fl);
    P->subsets.push_back(S);
    b0 = b;
  }
  const uint64_t cols = static_cast<uint64_t>(P->n_blocks) * kBlkCols;
  P->h_blocks = blocks;
  P->g_m.resize(G);
  for (uint32_t k = 0; k < G; ++k) P->g_m[k] = specs[k].m;
  if ((rc = dev_upload(ctx, &P->d_blocks, blocks, "source blocks")) ||
      (rc = dev_alloc(ctx, &P->d_sa, (P->state_floats = max_floats), "diffusion state")) ||
      (rc = dev_alloc(ctx, &P->d_sb, max_floats, "diffusion state")) ||
      (rc = dev_alloc(ctx, &P->d_active, cols, "column flags")) ||
      (rc = dev_alloc(ctx, &P->d_ell, cols, "column rounds")) ||
      (rc = dev_alloc(ctx, &P->d_colf, 2 * cols, "column bounds")) ||
      (rc = dev_alloc(ctx, &P->d_colu, 4 * cols + P->n_blocks, "column statistics")) ||
      (rc = dev_alloc(ctx, &P->d_dsum, cols, "column degree sums")))
    return fail(rc);
  P->per_cu_staged = static_cast<uint32_t>(block_post_per_cu(kPostLds));
  P->max_slots = static_cast<uint32_t>(static_cast<uint64_t>(ctx->num_cus) * P->per_cu_staged);
  P->slots = P->max_slots;
  TGTP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  *out = P;
  return TGTP_OK;
}

BlockArgs plan_args(const FastPlan* P, tgtp_dgraph* dg) {
  BlockArgs A{};
  const uint64_t cols = static_cast<uint64_t>(P->n_blocks) * kBlkCols;
  A.rowd = dg->rowd;
  A.adj = dg->adj;
  A.degf = dg->degf;
  A.taus = P->d_taus;
  A.preds_d = P->d_preds;
  A.nr_pos = P->d_nr;
  A.status = P->d_words + 2;
  A.g_eps = P->d_eps;
  A.g_logmd = P->d_logmd;
  A.g_gamma = P->d_gamma;
  A.blocks = P->d_blocks;
  A.xa = P->d_sa;
  A.xb = P->d_sb;
  A.active = P->d_active;
  A.ell = P->d_ell;
  A.gmax = P->d_colf;
  A.gmin = P->d_colf + cols;
  A.qsize = P->d_colu;
  A.st_mx = P->d_colu + cols;
  A.st_mn = P->d_colu + 2 * cols;
  A.st_q = P->d_colu + 3 * cols;
  A.last = P->d_colu + 4 * cols;
  A.st_dsum = P->d_dsum;
  A.next_src = P->d_words;
  A.light = P->light;
  A.small_bq = P->small_bq;
  A.tile_mode = P->tile_mode;
  A.round_form = P->round_form;
  A.adj_words = static_cast<uint32_t>(std::min<uint64_t>(dg->m, 0xffffffffull));
  A.hubs = P->d_hubs;
  A.hub_off = P->d_hub_off;
  return A;
}

TauFastArgs plan_tau_args(const FastPlan* P, tgtp_dgraph* dg) {
  TauFastArgs T{};
  T.idx = dg->idx;
  T.adj = dg->adj;
  T.und = P->d_und;
  T.tile_graph = P->d_tile_graph;
  T.graphs = P->d_tgraphs;
  T.n_tiles = P->n_tiles;
  T.X = dg->X;
  T.Xb = P->d_xb;
  T.Wt = P->d_wt;
  T.Wtb = P->d_wtb;
  T.n = static_cast<uint32_t>(dg->n);
  T.omega = P->omega;
  T.omega_pad = P->omega_pad;
  T.bounded = (P->mode & TGTP_TAU_REFERENCE) ? 0u : 1u;
  T.taus = P->d_taus;
  return T;
}

// The run itself, device work only, in three parts so that a capture can
// take each stream's share as a graph of its own: pre (clear, tau
// projection) on the main stream, the subsets (`parity` -1: all on `st`;
// 0 / 1: the even / odd ones, on the stream they run on), post
// (symmetrization) on the main stream.  Events 5 and 6 (when
// record_events) split tau / per-source / symmetrization.
int enqueue_pre(tgtp_ctx* ctx, FastPlan* P, tgtp_dgraph* dg, bool record_events) {
  int rc;
  if ((rc = launch_fill_words(ctx, P->d_words, 16)) ||
      (rc = launch_fill_i32(ctx, P->d_taus, P->m, 0)) ||
      (rc = launch_taus_fast(ctx, plan_tau_args(P, dg),
                             tgtp_mode_base(P->mode) == TGTP_MODE_FAST_BF16)))
    return rc;
  if (record_events) TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
  return TGTP_OK;
}

// Args of the main stream (second = false) or of the second stream, which
// has its own state buffers and post-kernel work counter.
BlockArgs stream_args(FastPlan* P, tgtp_dgraph* dg, const BlockArgs* override_args, bool second) {
  BlockArgs A = override_args ? *override_args : plan_args(P, dg);
  A.light = P->light;
  A.small_bq = P->small_bq;
  if (second) {
    A.xa = P->d_sa2;
    A.xb = P->d_sb2;
    A.next_src = P->d_words + 4;
  }
  return A;
}

// Subsets [0, count) whose index parity matches (`parity` -1: all; 0 / 1:
// even / odd), on the stream of their parity when `dual` (odd on the second
// stream), issued in subset order.
int enqueue_subsets(tgtp_ctx* ctx, FastPlan* P, tgtp_dgraph* dg, const BlockArgs* override_args,
                    int parity, bool dual, LaunchTimer* timer) {
  const BlockArgs A0 = stream_args(P, dg, override_args, false);
  const BlockArgs A1 = stream_args(P, dg, override_args, true);
  // TGTP_PROGRESS=1: an uncaptured run waits for each subset and reports it
  // on stderr (runs of minutes on large graphs)
  bool progress = env_int_batch("TGTP_PROGRESS", 0) != 0;
  if (progress) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    progress = hipStreamIsCapturing(ctx->stream, &cs) == hipSuccess &&
               cs == hipStreamCaptureStatusNone;
  }
  const double t_start = tgtp_now_ms();
  for (size_t i = 0; i < P->subsets.size(); ++i) {
    if (parity >= 0 && static_cast<int>(i & 1u) != parity) continue;
    const bool second = dual && (i & 1u);
    hipStream_t st = second ? P->side : ctx->stream;
    const BlockSubset& S = P->subsets[i];
    int rc = launch_block_subset(ctx, st, second ? A1 : A0, S,
                                 P->round_tiles.data() + S.round_tiles0,
                                 P->round_hubs.data() + S.round_tiles0,
                                 std::max(1u, P->slots * S.post_per_cu / P->per_cu_staged), timer);
    if (rc) return rc;
    if (progress) {
      TGTP_HIP_CHECK(ctx, hipStreamSynchronize(st));
      std::fprintf(stderr, "tgtp: source block subset %zu/%zu (%u rounds) done at %.1f s\n", i + 1,
                   P->subsets.size(), S.rounds, (tgtp_now_ms() - t_start) / 1e3);
      std::fflush(stderr);
    }
  }
  return TGTP_OK;
}

int enqueue_post(tgtp_ctx* ctx, FastPlan* P, const int32_t* src_of, bool record_events) {
  if (record_events) TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[6], ctx->stream));
  return launch_symmetrize_d(ctx, P->d_preds, src_of, P->m, P->d_out);
}

bool plan_dual(const FastPlan* P) { return P->dual && P->subsets.size() > 1; }

// Uncaptured run: odd subsets on the second stream (its own buffer pair),
// so one subset's post kernel (latency-bound walks) overlaps the next
// subset's rounds (bandwidth-bound).
int enqueue_run(tgtp_ctx* ctx, FastPlan* P, tgtp_dgraph* dg, const int32_t* src_of,
                bool record_events, const BlockArgs* override_args,
                LaunchTimer* timer = nullptr) {
  int rc;
  if ((rc = enqueue_pre(ctx, P, dg, record_events))) return rc;
  const bool dual = plan_dual(P);
  if (dual) {
    TGTP_HIP_CHECK(ctx, hipEventRecord(P->fork, ctx->stream));
    TGTP_HIP_CHECK(ctx, hipStreamWaitEvent(P->side, P->fork, 0));
  }
  if ((rc = enqueue_subsets(ctx, P, dg, override_args, -1, dual, timer))) return rc;
  if (dual) {
    TGTP_HIP_CHECK(ctx, hipEventRecord(P->join, P->side));
    TGTP_HIP_CHECK(ctx, hipStreamWaitEvent(ctx->stream, P->join, 0));
  }
  return enqueue_post(ctx, P, src_of, record_events);
}

// Second stream and state buffers for a plan with more than one subset
// (TGTP_BATCH_STREAMS=1 keeps one stream).
int enable_dual(tgtp_ctx* ctx, FastPlan* P) {
  if (P->subsets.size() < 2 || env_int_batch("TGTP_BATCH_STREAMS", 2) < 2) return TGTP_OK;
  int rc;
  if ((rc = dev_alloc(ctx, &P->d_sa2, P->state_floats, "diffusion state")) ||
      (rc = dev_alloc(ctx, &P->d_sb2, P->state_floats, "diffusion state")))
    return rc;
  TGTP_HIP_CHECK(ctx, hipStreamCreateWithFlags(&P->side, hipStreamNonBlocking));
  TGTP_HIP_CHECK(ctx, hipEventCreateWithFlags(&P->fork, hipEventDisableTiming));
  TGTP_HIP_CHECK(ctx, hipEventCreateWithFlags(&P->join, hipEventDisableTiming));
  P->dual = true;
  return TGTP_OK;
}

int timed_run(tgtp_ctx* ctx, FastPlan* P, tgtp_dgraph* dg, const int32_t* src_of, float* ms) {
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
  int rc = enqueue_run(ctx, P, dg, src_of, false, nullptr);
  if (rc) return rc;
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[7], ctx->stream));
  TGTP_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev[7]));
  TGTP_HIP_CHECK(ctx, hipEventElapsedTime(ms, ctx->ev[4], ctx->ev[7]));
  return TGTP_OK;
}

int autotune(tgtp_ctx* ctx, FastPlan* P, tgtp_dgraph* dg, const int32_t* src_of) {
  float best = 0;
  int rc = timed_run(ctx, P, dg, src_of, &best);
  if (rc) return rc;
  auto trial = [&](uint32_t* knob, std::initializer_list<uint32_t> values) -> int {
    const uint32_t keep = *knob;
    uint32_t pick = keep;
    for (uint32_t v : values) {
      if (v == keep) continue;
      *knob = v;
      float ms = 0;
      int r = timed_run(ctx, P, dg, src_of, &ms);
      if (r) return r;
      if (ms < best) {
        best = ms;
        pick = v;
      }
    }
    *knob = pick;
    return TGTP_OK;
  };
  if ((rc = trial(&P->light, {32, 64}))) return rc;
  if ((rc = trial(&P->small_bq, {8, 32}))) return rc;
  const uint32_t full = P->max_slots;
  if ((rc = trial(&P->slots, {full, std::max(1u, (3 * full) / 4)}))) return rc;
  return TGTP_OK;
}

// Captures one part of the run on `st` into P->graph[k] / P->exec[k].
template <typename F>
int capture_part(tgtp_ctx* ctx, FastPlan* P, int k, hipStream_t st, F&& enqueue) {
  TGTP_HIP_CHECK(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue();
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return set_error(ctx, TGTP_E_HIP, "hipStreamEndCapture failed");
  P->graph[k] = g;
  if (hipGraphInstantiate(&P->exec[k], g, nullptr, nullptr, 0) != hipSuccess) {
    P->exec[k] = nullptr;
    return set_error(ctx, TGTP_E_HIP, "hipGraphInstantiate failed");
  }
  return TGTP_OK;
}

// A single-stream plan is one graph.  A two-stream plan is captured as one
// graph per stream share -- pre (main), even subsets (main), odd subsets
// (second stream), post (main) -- replayed on their streams with the fork /
// join events between them: one graph holding both streams' branches ran
// its branches without their overlap (6% slower than the uncaptured run on
// the 256-graph batch, while a one-stream graph matched its uncaptured run).
int capture(tgtp_ctx* ctx, FastPlan* P, tgtp_dgraph* dg, const int32_t* src_of) {
  if (!plan_dual(P))
    return capture_part(ctx, P, 0, ctx->stream,
                        [&] { return enqueue_run(ctx, P, dg, src_of, false, nullptr); });
  int rc;
  if ((rc = capture_part(ctx, P, 0, ctx->stream, [&] { return enqueue_pre(ctx, P, dg, false); })) ||
      (rc = capture_part(ctx, P, 1, ctx->stream,
                         [&] { return enqueue_subsets(ctx, P, dg, nullptr, 0, true, nullptr); })) ||
      (rc = capture_part(ctx, P, 2, P->side,
                         [&] { return enqueue_subsets(ctx, P, dg, nullptr, 1, true, nullptr); })) ||
      (rc = capture_part(ctx, P, 3, ctx->stream,
                         [&] { return enqueue_post(ctx, P, src_of, false); })))
    return rc;
  return TGTP_OK;
}

// Replays the captured run on the context's stream (and the second stream).
int launch_captured(tgtp_ctx* ctx, FastPlan* P) {
  hipStream_t s = ctx->stream;
  if (!P->exec[1]) {
    TGTP_HIP_CHECK(ctx, hipGraphLaunch(P->exec[0], s));
    return TGTP_OK;
  }
  TGTP_HIP_CHECK(ctx, hipGraphLaunch(P->exec[0], s));
  TGTP_HIP_CHECK(ctx, hipEventRecord(P->fork, s));
  TGTP_HIP_CHECK(ctx, hipStreamWaitEvent(P->side, P->fork, 0));
  TGTP_HIP_CHECK(ctx, hipGraphLaunch(P->exec[2], P->side));
  TGTP_HIP_CHECK(ctx, hipEventRecord(P->join, P->side));
  TGTP_HIP_CHECK(ctx, hipGraphLaunch(P->exec[1], s));
  TGTP_HIP_CHECK(ctx, hipStreamWaitEvent(s, P->join, 0));
  TGTP_HIP_CHECK(ctx, hipGraphLaunch(P->exec[3], s));
  return TGTP_OK;
}

int check_status(tgtp_ctx* ctx, FastPlan* P) {
  uint32_t words[3] = {0, 0, 0};
  TGTP_HIP_CHECK(ctx, hipMemcpy(words, P->d_words, 12, hipMemcpyDeviceToHost));
  if (words[2] & 1u)
    return set_error(ctx, TGTP_E_UNSUPPORTED,
                     "a walk count exceeded 2^36 walks per edge; scores are not computed");
  if (words[2] & 2u)
    return set_error(ctx, TGTP_E_HIP, "a source diffused beyond its round bound");
  return TGTP_OK;
}

bool plan_matches(const FastPlan* P, const tgtp_dgraph* dg, const tgtp_params* p) {
  return P && P->eig_gen == dg->eig_gen && P->eps == p->epsilon && P->delta == p->delta &&
         P->gamma == p->gamma && P->mode == p->mode && P->omega == dg->eig_omega;
}

int single_plan(tgtp_ctx* ctx, tgtp_dgraph* dg, const tgtp_eigens* e, const tgtp_params* p,
                FastPlan** out) {
  auto* prep = static_cast<TgtpPrep*>(dg->prep);
  auto* cached = static_cast<FastPlan*>(prep->fast_plan);
  if (plan_matches(cached, dg, p)) {
    *out = cached;
    return TGTP_OK;
  }
  if (cached) free_plan(cached);
  prep->fast_plan = nullptr;
  std::vector<GraphSpec> specs(1);
  specs[0] = GraphSpec{0u, static_cast<uint32_t>(dg->n), 0ull, dg->m, e->values, p->epsilon,
                       p->delta, p->gamma};
  FastPlan* P = nullptr;
  int rc = build_plan(ctx, dg, specs, e->omega, p->mode, &P);
  if (rc) return rc;
  // a graph whose blocks take several subsets alternates them between two
  // streams as a batch does
  if ((rc = enable_dual(ctx, P))) {
    free_plan(P);
    return rc;
  }
  prep->fast_plan = P;
  prep->fast_plan_free = [](void* q) { free_plan(static_cast<FastPlan*>(q)); };
  *out = P;
  return TGTP_OK;
}

}  // namespace

// ---- single graph -------------------------------------------------------------

int tgtp_run_score_fast(tgtp_ctx* ctx, tgtp_dgraph* dg, const tgtp_eigens* e,
                        const tgtp_params* p, float* scores_out, tgtp_stats* stats,
                        const TraceOut* tr) {
  const double t0 = tgtp_now_ms();
  if (e->omega > 128 || e->omega < 3)
    return set_error(ctx, TGTP_E_UNSUPPORTED, "fast modes need 3 <= omega <= 128");
  auto* prep = static_cast<TgtpPrep*>(dg->prep);
  FastPlan* P = nullptr;
  int rc = single_plan(ctx, dg, e, p, &P);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  float* tr_buf = nullptr;
  BlockArgs A = plan_args(P, dg);
  if (tr) {
    const uint64_t n = dg->n, m = dg->m;
    tr_buf = static_cast<float*>(ctx_buffer(ctx, kBufTrace, (5 * n + m) * 4 + 64));
    if (!tr_buf) return set_error(ctx, TGTP_E_NOMEM, "trace buffers");
    A.tr_ell = reinterpret_cast<int32_t*>(tr_buf);
    A.tr_emax = tr_buf + n;
    A.tr_gmax = tr_buf + 2 * n;
    A.tr_gmin = tr_buf + 3 * n;
    A.tr_qsize = reinterpret_cast<uint32_t*>(tr_buf + 4 * n);
    A.tr_x = tr_buf + 5 * n;
    if ((rc = launch_fill_i32(ctx, reinterpret_cast<int32_t*>(tr_buf), 5 * n + m, 0))) return rc;
  }
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[4], s));
  if ((rc = enqueue_run(ctx, P, dg, prep->d_src_of, true, &A))) return rc;
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[7], s));
  TGTP_HIP_CHECK(ctx, hipMemcpyAsync(scores_out, P->d_out, dg->m * 4, hipMemcpyDeviceToHost, s));
  if (tr) {
    const uint64_t n = dg->n, m = dg->m;
    auto d2h = [&](void* dst, const void* src, uint64_t bytes) -> int {
      if (dst) TGTP_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
      return TGTP_OK;
    };
    if ((rc = d2h(tr->taus, P->d_taus, m * 4)) || (rc = d2h(tr->nr, P->d_nr, m * 8)) ||
        (rc = d2h(tr->walk_sum, A.tr_x, m * 4)) || (rc = d2h(tr->ell, A.tr_ell, n * 4)) ||
        (rc = d2h(tr->edge_max, A.tr_emax, n * 4)) || (rc = d2h(tr->gmax, A.tr_gmax, n * 4)) ||
        (rc = d2h(tr->gmin, A.tr_gmin, n * 4)) || (rc = d2h(tr->q_size, A.tr_qsize, n * 4)))
      return rc;
  }
  TGTP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  if (tr && tr->preds_row) {
    std::vector<double> pd(dg->m);
    TGTP_HIP_CHECK(ctx, hipMemcpy(pd.data(), P->d_preds, dg->m * 8, hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < dg->m; ++i) tr->preds_row[i] = static_cast<float>(pd[i]);
  }
  if ((rc = check_status(ctx, P))) return rc;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    float a = 0, b = 0, c = 0;
    TGTP_HIP_CHECK(ctx, hipEventElapsedTime(&a, ctx->ev[4], ctx->ev[7]));
    TGTP_HIP_CHECK(ctx, hipEventElapsedTime(&b, ctx->ev[4], ctx->ev[5]));
    TGTP_HIP_CHECK(ctx, hipEventElapsedTime(&c, ctx->ev[5], ctx->ev[6]));
    stats->kernel_ms = a;
    stats->taus_ms = b;
    stats->diffuse_ms = c;
    stats->e2e_ms = tgtp_now_ms() - t0;
  }
  return TGTP_OK;
}

int tgtp_run_taus_fast(tgtp_ctx* ctx, tgtp_dgraph* dg, const tgtp_eigens* e, float eps_half,
                       int32_t mode, int32_t* d_taus) {
  if (e->omega > 128 || e->omega < 3)
    return set_error(ctx, TGTP_E_UNSUPPORTED, "fast modes need 3 <= omega <= 128");
  tgtp_params p{static_cast<double>(eps_half) * 2.0, 1.0 / static_cast<double>(dg->n), 10, mode};
  FastPlan* P = nullptr;
  int rc = single_plan(ctx, dg, e, &p, &P);
  if (rc) return rc;
  // float(p.epsilon / 2) == eps_half exactly (doubling and halving are exact)
  if ((rc = launch_fill_i32(ctx, P->d_taus, P->m, 0)) ||
      (rc = launch_taus_fast(ctx, plan_tau_args(P, dg),
                             tgtp_mode_base(mode) == TGTP_MODE_FAST_BF16)))
    return rc;
  TGTP_HIP_CHECK(ctx, hipMemcpyAsync(d_taus, P->d_taus, dg->m * 4, hipMemcpyDeviceToDevice,
                                     ctx->stream));
  return TGTP_OK;
}

// Source subset: a plan whose blocks cover only `ranges` (not cached on the
// dgraph), one run, every score and row copied back (the caller keeps the
// positions it can use).
int tgtp_run_score_fast_sources(tgtp_ctx* ctx, tgtp_dgraph* dg, const tgtp_eigens* e,
                                const tgtp_params* p,
                                const std::vector<std::pair<uint32_t, uint32_t>>& ranges,
                                float* scores_all, float* rows_all, tgtp_stats* stats) {
  const double t0 = tgtp_now_ms();
  if (e->omega > 128 || e->omega < 3)
    return set_error(ctx, TGTP_E_UNSUPPORTED, "fast modes need 3 <= omega <= 128");
  auto* prep = static_cast<TgtpPrep*>(dg->prep);
  std::vector<GraphSpec> specs(1);
  specs[0] = GraphSpec{0u, static_cast<uint32_t>(dg->n), 0ull, dg->m, e->values, p->epsilon,
                       p->delta, p->gamma};
  FastPlan* P = nullptr;
  int rc = build_plan(ctx, dg, specs, e->omega, p->mode, &P, &ranges);
  if (rc) return rc;
  auto done = [&](int r) {
    free_plan(P);
    return r;
  };
  if ((rc = enable_dual(ctx, P))) return done(rc);
  hipStream_t s = ctx->stream;
  if (hipEventRecord(ctx->ev[4], s) != hipSuccess) return done(set_error(ctx, TGTP_E_HIP, "event"));
  if ((rc = enqueue_run(ctx, P, dg, prep->d_src_of, true, nullptr))) return done(rc);
  std::vector<double> pd(dg->m);
  if (hipEventRecord(ctx->ev[7], s) != hipSuccess ||
      hipMemcpyAsync(scores_all, P->d_out, dg->m * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(pd.data(), P->d_preds, dg->m * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return done(set_error(ctx, TGTP_E_HIP, "subset run"));
  for (uint64_t i = 0; i < dg->m; ++i) rows_all[i] = static_cast<float>(pd[i]);
  if ((rc = check_status(ctx, P))) return done(rc);
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    float a = 0, b = 0, c = 0;
    (void)hipEventElapsedTime(&a, ctx->ev[4], ctx->ev[7]);
    (void)hipEventElapsedTime(&b, ctx->ev[4], ctx->ev[5]);
    (void)hipEventElapsedTime(&c, ctx->ev[5], ctx->ev[6]);
    stats->kernel_ms = a;
    stats->taus_ms = b;
    stats->diffuse_ms = c;
    stats->work_items = P->n_sources;
    stats->e2e_ms = tgtp_now_ms() - t0;
  }
  return done(TGTP_OK);
}

// ---- batched C ABI ---------------------------------------------------------------

struct tgtp_batch {
  int32_t count = 0;
  int32_t mode = 0;
  int32_t flags = 0;
  tgtp_dgraph* dg = nullptr;          // the union (FAST modes)
  FastPlan* plan = nullptr;
  std::vector<uint64_t> eoff;         // CSR offset of each graph in the union
  std::vector<uint64_t> m;            // CSR entries of each graph
  // FAITHFUL mode: graphs run one after another
  std::vector<tgtp_dgraph*> singles;
  std::vector<tgtp_eigens> eig;
  std::vector<tgtp_params> prm;
  float tune_ms = 0;
};

extern "C" void tgtp_batch_destroy(tgtp_ctx* ctx, tgtp_batch* b) {
  if (!b) return;
  if (b->plan) free_plan(b->plan);
  if (b->dg) tgtp_graph_free(ctx, b->dg);
  for (auto* d : b->singles) tgtp_graph_free(ctx, d);
  delete b;
}

extern "C" int tgtp_batch_create(tgtp_ctx* ctx, int32_t count, const tgtp_graph* graphs,
                                 const tgtp_eigens* eigens, const tgtp_params* params,
                                 int32_t flags, tgtp_batch** out) {
  if (!ctx || !out) return TGTP_E_ARG;
  *out = nullptr;
  if (count < 1 || !graphs || !eigens || !params)
    return set_error(ctx, TGTP_E_ARG, "empty batch or null arrays");
  const int32_t mode = params[0].mode;
  const int32_t omega = eigens[0].omega;
  for (int32_t k = 0; k < count; ++k) {
    const tgtp_params& p = params[k];
    if (!(p.epsilon > 0 && p.epsilon < 1) || p.gamma < 1 || !(p.delta > 0))
      return set_error(ctx, TGTP_E_ARG,
                       "require 0 < epsilon < 1, integer gamma >= 1, delta > 0 for every graph");
    if (p.mode != mode) return set_error(ctx, TGTP_E_ARG, "all graphs of a batch need one mode");
    if (!tgtp_mode_valid(p.mode)) return set_error(ctx, TGTP_E_ARG, "unknown tgtp mode");
    if (!graphs[k].idx || graphs[k].n == 0 || !eigens[k].values || !eigens[k].vectors ||
        eigens[k].omega < 2 || static_cast<uint64_t>(eigens[k].omega) > graphs[k].n)
      return set_error(ctx, TGTP_E_ARG, "graph or eigenpairs missing / omega out of range");
    // each graph on its own: the union's checks would not see an id that
    // points into the next graph's vertices
    const tgtp_graph& g = graphs[k];
    if ((g.m && !g.adj) || g.idx[0] != 0 || g.idx[g.n] != g.m)
      return set_error(ctx, TGTP_E_ARG, "graph " + std::to_string(k) +
                                            ": null adjacency or idx[0] != 0 or idx[n] != m");
    for (uint64_t p = 0; p < g.m; ++p)
      if (g.adj[p] >= g.n)
        return set_error(ctx, TGTP_E_ARG, "graph " + std::to_string(k) + ": neighbour id >= n");
    for (uint64_t v = 0; v < graphs[k].n; ++v)
      if (graphs[k].idx[v + 1] == graphs[k].idx[v])
        return set_error(ctx, TGTP_E_ARG,
                         "graph vertex IDs must be contiguous with no isolated vertices");
  }
  TGTP_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  auto* b = new (std::nothrow) tgtp_batch();
  if (!b) return set_error(ctx, TGTP_E_NOMEM, "host allocation");
  b->count = count;
  b->mode = mode;
  b->flags = flags;
  auto fail = [&](int rc) {
    tgtp_batch_destroy(ctx, b);
    return rc;
  };
  int rc;
  if (tgtp_mode_base(mode) == TGTP_MODE_FAITHFUL) {
    for (int32_t k = 0; k < count; ++k) {
      tgtp_dgraph* d = nullptr;
      if ((rc = tgtp_graph_upload(ctx, &graphs[k], &d))) return fail(rc);
      b->singles.push_back(d);
      b->eig.push_back(eigens[k]);
      b->prm.push_back(params[k]);
      b->m.push_back(graphs[k].m);
    }
    *out = b;
    return TGTP_OK;
  }
  for (int32_t k = 0; k < count; ++k)
    if (eigens[k].omega != omega)
      return fail(set_error(ctx, TGTP_E_UNSUPPORTED, "fast batches need one omega for all graphs"));
  if (omega < 3 || omega > 128)
    return fail(set_error(ctx, TGTP_E_UNSUPPORTED, "fast modes need 3 <= omega <= 128"));
  // disjoint union on the host
  uint64_t N = 0, M = 0;
  for (int32_t k = 0; k < count; ++k) {
    N += graphs[k].n;
    M += graphs[k].m;
  }
  if (N > 0xFFFFFFFEull || M > 0xFFFFFFFFull)
    return fail(set_error(ctx, TGTP_E_UNSUPPORTED, "batch exceeds 32-bit CSR capacity"));
  std::vector<uint32_t> uidx(N + 1), uadj(M);
  std::vector<float> ux(N * static_cast<uint64_t>(omega));
  std::vector<GraphSpec> specs(count);
  uint64_t vo = 0, eo = 0;
  for (int32_t k = 0; k < count; ++k) {
    const tgtp_graph& g = graphs[k];
    for (uint64_t v = 0; v < g.n; ++v) uidx[vo + v] = static_cast<uint32_t>(eo + g.idx[v]);
    for (uint64_t p = 0; p < g.m; ++p) uadj[eo + p] = static_cast<uint32_t>(vo + g.adj[p]);
    std::memcpy(ux.data() + vo * omega, eigens[k].vectors, g.n * omega * sizeof(float));
    specs[k] = GraphSpec{static_cast<uint32_t>(vo), static_cast<uint32_t>(g.n), eo, g.m,
                         eigens[k].values, params[k].epsilon, params[k].delta, params[k].gamma};
    b->eoff.push_back(eo);
    b->m.push_back(g.m);
    vo += g.n;
    eo += g.m;
  }
  uidx[N] = static_cast<uint32_t>(M);
  const tgtp_graph ug{N, M, uidx.data(), uadj.data()};
  if ((rc = tgtp_graph_upload(ctx, &ug, &b->dg))) return fail(rc);
  if ((rc = tgtp_ensure_prep(ctx, b->dg))) return fail(rc);
  if ((rc = dev_alloc(ctx, &b->dg->X, ux.size(), "eigenvectors"))) return fail(rc);
  TGTP_HIP_CHECK(ctx, hipMemcpyAsync(b->dg->X, ux.data(), ux.size() * 4, hipMemcpyHostToDevice,
                                     ctx->stream));
  b->dg->eig_omega = omega;
  b->dg->eig_key = b;  // never matches a caller pointer
  if ((rc = build_plan(ctx, b->dg, specs, omega, mode, &b->plan))) return fail(rc);
  if ((rc = enable_dual(ctx, b->plan))) return fail(rc);
  auto* prep = static_cast<TgtpPrep*>(b->dg->prep);
  if (flags & TGTP_BATCH_AUTOTUNE) {
    const double t0 = tgtp_now_ms();
    if ((rc = autotune(ctx, b->plan, b->dg, prep->d_src_of))) return fail(rc);
    b->tune_ms = static_cast<float>(tgtp_now_ms() - t0);
  }
  if (flags & TGTP_BATCH_GRAPH)
    if ((rc = capture(ctx, b->plan, b->dg, prep->d_src_of))) return fail(rc);
  *out = b;
  return TGTP_OK;
}

extern "C" int tgtp_batch_run(tgtp_ctx* ctx, tgtp_batch* b, float* const* scores_out,
                              tgtp_stats* stats) {
  if (!ctx) return TGTP_E_ARG;
  if (!b || !scores_out) return set_error(ctx, TGTP_E_ARG, "null batch or outputs");
  const double t0 = tgtp_now_ms();
  TGTP_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  int rc;
  if (tgtp_mode_base(b->mode) == TGTP_MODE_FAITHFUL) {
    double kms = 0;
    for (int32_t k = 0; k < b->count; ++k) {
      tgtp_stats st{};
      if ((rc = tgtp_score_dev(ctx, b->singles[k], &b->eig[k], &b->prm[k], scores_out[k], &st)))
        return rc;
      kms += st.kernel_ms;
    }
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->kernel_ms = kms;
      stats->e2e_ms = tgtp_now_ms() - t0;
    }
    return TGTP_OK;
  }
  hipStream_t s = ctx->stream;
  auto* prep = static_cast<TgtpPrep*>(b->dg->prep);
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[4], s));
  if (b->plan->exec[0]) {
    if ((rc = launch_captured(ctx, b->plan))) return rc;
  } else if ((rc = enqueue_run(ctx, b->plan, b->dg, prep->d_src_of, false, nullptr))) {
    return rc;
  }
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[7], s));
  for (int32_t k = 0; k < b->count; ++k)
    if (scores_out[k] && b->m[k])
      TGTP_HIP_CHECK(ctx, hipMemcpyAsync(scores_out[k], b->plan->d_out + b->eoff[k], b->m[k] * 4,
                                         hipMemcpyDeviceToHost, s));
  TGTP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  if ((rc = check_status(ctx, b->plan))) return rc;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    float a = 0;
    TGTP_HIP_CHECK(ctx, hipEventElapsedTime(&a, ctx->ev[4], ctx->ev[7]));
    stats->kernel_ms = a;
    stats->e2e_ms = tgtp_now_ms() - t0;
    stats->work_items = b->plan->n_sources;
    stats->taus_ms = b->tune_ms;
  }
  return TGTP_OK;
}

extern "C" int tgtp_batch_profile(tgtp_ctx* ctx, tgtp_batch* b, tgtp_profile* out) {
  if (!ctx) return TGTP_E_ARG;
  if (!b || !out) return set_error(ctx, TGTP_E_ARG, "null batch or output");
  if (!b->plan) return set_error(ctx, TGTP_E_UNSUPPORTED, "profiles cover FAST batches only");
  TGTP_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  FastPlan* P = b->plan;
  auto* prep = static_cast<TgtpPrep*>(b->dg->prep);
  hipStream_t s = ctx->stream;
  LaunchTimer T;
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[4], s));
  int rc = enqueue_run(ctx, P, b->dg, prep->d_src_of, true, nullptr, &T);
  if (rc) return rc;
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[7], s));
  TGTP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  if (P->side) TGTP_HIP_CHECK(ctx, hipStreamSynchronize(P->side));
  if ((rc = check_status(ctx, P))) return rc;
  std::memset(out, 0, sizeof(*out));
  float a = 0;
  TGTP_HIP_CHECK(ctx, hipEventElapsedTime(&a, ctx->ev[4], ctx->ev[7]));
  out->total_ms = a;
  TGTP_HIP_CHECK(ctx, hipEventElapsedTime(&a, ctx->ev[4], ctx->ev[5]));
  out->taus_ms = a;
  double ms[LaunchTimer::kKinds] = {0, 0, 0};
  uint64_t cnt[LaunchTimer::kKinds] = {0, 0, 0};
  for (size_t i = 0; i < T.kind.size(); ++i) {
    float e = 0;
    TGTP_HIP_CHECK(ctx, hipEventElapsedTime(&e, T.ev[2 * i], T.ev[2 * i + 1]));
    ms[T.kind[i]] += e;
    ++cnt[T.kind[i]];
  }
  out->round_ms = ms[LaunchTimer::kRound];
  out->round_launches = cnt[LaunchTimer::kRound];
  out->stop_ms = ms[LaunchTimer::kStop];
  out->stop_launches = cnt[LaunchTimer::kStop];
  out->post_ms = ms[LaunchTimer::kPost];
  out->post_launches = cnt[LaunchTimer::kPost];
  // rounds each source (ell) and each block (last) took
  const uint64_t cols = static_cast<uint64_t>(P->n_blocks) * kBlkCols;
  std::vector<int32_t> ell(cols);
  std::vector<uint32_t> last(P->n_blocks);
  TGTP_HIP_CHECK(ctx, hipMemcpy(ell.data(), P->d_ell, cols * 4, hipMemcpyDeviceToHost));
  TGTP_HIP_CHECK(ctx, hipMemcpy(last.data(), P->d_colu + 4 * cols, P->n_blocks * 4ull,
                                hipMemcpyDeviceToHost));
  uint64_t bytes = 0, src_rounds = 0;
  for (uint32_t k = 0; k < P->n_blocks; ++k) {
    const BlockInfo& B = P->h_blocks[k];
    const uint64_t mg = P->g_m[B.gi], ng = B.n;
    if (last[k] > 1) bytes += (last[k] - 1ull) * (4 * mg + 8 * ng);  // ids + row descriptors
    for (uint32_t j = 0; j < B.cnt; ++j) {
      const int32_t e = ell[static_cast<uint64_t>(k) * kBlkCols + j];
      src_rounds += static_cast<uint64_t>(std::max(e, 0));
      if (e > 1) bytes += (e - 1ull) * 4 * (mg + ng);  // its value per edge read, per row written
    }
  }
  out->round_bytes = bytes;
  out->source_rounds = src_rounds;
  return TGTP_OK;
}

extern "C" int tgtp_batch_config(const tgtp_batch* b, uint32_t* light, uint32_t* small_bq,
                                 uint32_t* slots, uint32_t* max_slots) {
  if (!b || !b->plan) return TGTP_E_ARG;
  if (light) *light = b->plan->light;
  if (small_bq) *small_bq = b->plan->small_bq;
  if (slots) *slots = b->plan->slots;
  if (max_slots) *max_slots = b->plan->max_slots;
  return TGTP_OK;
}

extern "C" int tgtp_score_batch(tgtp_ctx* ctx, int32_t count, const tgtp_graph* graphs,
                                const tgtp_eigens* eigens, const tgtp_params* params,
                                float* const* scores_out, tgtp_stats* stats) {
  tgtp_batch* b = nullptr;
  int rc = tgtp_batch_create(ctx, count, graphs, eigens, params, 0, &b);
  if (rc) return rc;
  rc = tgtp_batch_run(ctx, b, scores_out, stats);
  tgtp_batch_destroy(ctx, b);
  return rc;
}
