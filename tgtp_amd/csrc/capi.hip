// capi.hip -- the C ABI of include/tgtp_hip.h: context, device graph,
// argument validation and dispatch to the kernels.
#include <chrono>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "common.hpp"

// Scratch of tgtp_mt_walks_noalloc: the walk counters of at most
// kNoAllocMaxN vertices, then the two 64-bit sums.  Every GPU the library
// runs on holds its own copy (16 MiB) until the process ends.  In host code
// `scratch` is only a stand-in for the device variable: its address on each
// device comes from hipGetSymbolAddress, once per device, in
// tgtp_ctx_create.
constexpr uint64_t kNoAllocMaxN = 1ull << 22;
static __device__ __align__(256) uint32_t scratch[kNoAllocMaxN + 64];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// This device's address of `scratch` (the device must be current); null if
// the runtime cannot resolve it.
void* device_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

__global__ void build_rowd_kernel(const uint32_t* __restrict__ idx, uint64_t n,
                                  uint2* __restrict__ rowd, float* __restrict__ degf) {
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t a = idx[v], b = idx[v + 1];
    rowd[v] = make_uint2(a, b - a);
    degf[v] = static_cast<float>(b - a);
  }
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

int check_host_graph(tgtp_ctx* ctx, const tgtp_graph* g) {
  if (!g || !g->idx || (g->m && !g->adj) || g->n == 0)
    return set_error(ctx, TGTP_E_ARG, "graph: null arrays or n == 0");
  if (g->n > 0xFFFFFFFEull || g->m > 0xFFFFFFFFull)
    return set_error(ctx, TGTP_E_ARG, "graph exceeds CSR index capacity");
  if (g->idx[0] != 0 || g->idx[g->n] != g->m)
    return set_error(ctx, TGTP_E_ARG, "graph: idx[0] != 0 or idx[n] != m");
  // every kernel indexes rows and neighbours without bounds checks: make sure
  // rows are monotone and neighbour ids are vertices
  for (uint64_t v = 0; v < g->n; ++v)
    if (g->idx[v + 1] < g->idx[v])
      return set_error(ctx, TGTP_E_ARG, "graph: idx is not non-decreasing");
  uint32_t bad = 0;
  for (uint64_t p = 0; p < g->m; ++p) bad |= (g->adj[p] >= g->n);
  if (bad) return set_error(ctx, TGTP_E_ARG, "graph: neighbour id >= n");
  return TGTP_OK;
}

}  // namespace

int set_error(tgtp_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

void* ctx_buffer(tgtp_ctx* ctx, int slot, size_t bytes) {
  DevBuf& b = ctx->bufs[slot];
  if (b.bytes >= bytes && b.p) return b.p;
  if (b.p) {
    (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes < 256 ? 256 : bytes;
  if (hipMalloc(&b.p, want) != hipSuccess) {
    b.p = nullptr;
    return nullptr;
  }
  b.bytes = want;
  return b.p;
}

extern "C" const char* tgtp_strerror(int code) {
  switch (code) {
    case TGTP_OK: return "ok";
    case TGTP_E_ARG: return "invalid argument";
    case TGTP_E_HIP: return "HIP runtime error";
    case TGTP_E_NOMEM: return "out of memory";
    case TGTP_E_IO: return "I/O error";
    case TGTP_E_UNSUPPORTED: return "unsupported shape";
    default: return "unknown error";
  }
}

extern "C" int tgtp_ctx_create(int device, tgtp_ctx** out) {
  if (!out) return TGTP_E_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return TGTP_E_HIP;
  if (device < 0 || device >= count) return TGTP_E_ARG;
  tgtp_ctx* ctx = new (std::nothrow) tgtp_ctx();
  if (!ctx) return TGTP_E_NOMEM;
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete ctx;
    return TGTP_E_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete ctx;
    return TGTP_E_HIP;
  }
  ctx->num_cus = prop.multiProcessorCount;
  ctx->scratch = device_scratch(device);  // checked by the entry point that uses it
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return TGTP_E_HIP;
  }
  for (auto& e : ctx->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      tgtp_ctx_destroy(ctx);
      return TGTP_E_HIP;
    }
  }
  *out = ctx;
  return TGTP_OK;
}

extern "C" void tgtp_ctx_destroy(tgtp_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& b : ctx->bufs)
    if (b.p) (void)hipFree(b.p);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" const char* tgtp_ctx_last_error(const tgtp_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int tgtp_graph_upload(tgtp_ctx* ctx, const tgtp_graph* g, tgtp_dgraph** out) {
  if (!ctx || !out) return TGTP_E_ARG;
  *out = nullptr;
  int rc = check_host_graph(ctx, g);
  if (rc) return rc;
  TGTP_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  tgtp_dgraph* dg = new (std::nothrow) tgtp_dgraph();
  if (!dg) return set_error(ctx, TGTP_E_NOMEM, "host allocation");
  dg->device = ctx->device;
  dg->n = g->n;
  dg->m = g->m;
  for (uint64_t v = 0; v < g->n; ++v)
    dg->max_deg = std::max<uint64_t>(dg->max_deg, g->idx[v + 1] - g->idx[v]);
  auto fail = [&](const char* what) {
    tgtp_graph_free(ctx, dg);
    return set_error(ctx, TGTP_E_NOMEM, std::string("device allocation: ") + what);
  };
  if (hipMalloc(&dg->idx, (g->n + 1) * sizeof(uint32_t)) != hipSuccess) return fail("idx");
  if (hipMalloc(&dg->adj, (g->m ? g->m : 1) * sizeof(uint32_t)) != hipSuccess) return fail("adj");
  if (hipMalloc(&dg->rowd, g->n * sizeof(uint2)) != hipSuccess) return fail("rowd");
  if (hipMalloc(&dg->degf, g->n * sizeof(float)) != hipSuccess) return fail("degf");
  hipStream_t s = ctx->stream;
  TGTP_HIP_CHECK(ctx, hipMemcpyAsync(dg->idx, g->idx, (g->n + 1) * sizeof(uint32_t),
                                     hipMemcpyHostToDevice, s));
  if (g->m)
    TGTP_HIP_CHECK(ctx, hipMemcpyAsync(dg->adj, g->adj, g->m * sizeof(uint32_t),
                                       hipMemcpyHostToDevice, s));
  const uint64_t blocks = std::min<uint64_t>(4096, (g->n + 255) / 256);
  build_rowd_kernel<<<static_cast<uint32_t>(blocks), 256, 0, s>>>(dg->idx, g->n, dg->rowd, dg->degf);
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  TGTP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  *out = dg;
  return TGTP_OK;
}

extern "C" void tgtp_graph_free(tgtp_ctx* ctx, tgtp_dgraph* dg) {
  if (!dg) return;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  (void)hipSetDevice(dg->device);
  if (dg->prep && dg->prep_free) dg->prep_free(dg->prep);
  void* ptrs[] = {dg->idx, dg->adj, dg->rowd, dg->rev, dg->degf, dg->lam, dg->X, dg->walk_adj,
                   dg->part_dev, dg->part_adj};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete dg;
}

// Shared body of tgtp_mt_walks and tgtp_mt_walks_noalloc: validation, then
// the walks with the given device counters [n] and sums [2].
static int mt_walks_check(tgtp_ctx* ctx, const tgtp_dgraph* dg, uint32_t src_begin,
                          uint32_t src_end, uint32_t n_walks, int32_t variant) {
  if (!dg) return set_error(ctx, TGTP_E_ARG, "null device graph");
  if (variant != TGTP_WALK_MT_REGULAR && variant != TGTP_WALK_MT_SIMD)
    return set_error(ctx, TGTP_E_ARG, "unknown walk variant");
  if (variant == TGTP_WALK_MT_SIMD && n_walks % 8 != 0)
    return set_error(ctx, TGTP_E_ARG, "n_walks must be a multiple of 8");
  if (src_end > dg->n || src_begin > src_end)
    return set_error(ctx, TGTP_E_ARG, "source range outside [0, n]");
  return TGTP_OK;
}

static int mt_walks_run(tgtp_ctx* ctx, const tgtp_dgraph* dg, uint32_t src_begin, uint32_t src_end,
                        uint32_t depth, uint32_t n_walks, int32_t variant, uint32_t* counters_out,
                        uint64_t* visit_count, uint64_t* checksum, tgtp_stats* stats,
                        uint32_t* d_counters, unsigned long long* d_sums, double t0,
                        bool may_alloc) {
  float walk_ms = 0;
  int rc = launch_mt_walks(ctx, dg, src_begin, src_end, depth, n_walks, variant, d_counters,
                           d_sums, &walk_ms, may_alloc);
  if (rc) return rc;
  unsigned long long sums[2] = {0, 0};
  TGTP_HIP_CHECK(ctx, hipMemcpyAsync(sums, d_sums, sizeof(sums), hipMemcpyDeviceToHost,
                                     ctx->stream));
  if (counters_out)
    TGTP_HIP_CHECK(ctx, hipMemcpyAsync(counters_out, d_counters, dg->n * sizeof(uint32_t),
                                       hipMemcpyDeviceToHost, ctx->stream));
  TGTP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (visit_count) *visit_count = sums[0];
  if (checksum) *checksum = sums[1];
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    float all_ms = 0;
    TGTP_HIP_CHECK(ctx, hipEventElapsedTime(&all_ms, ctx->ev[0], ctx->ev[3]));
    stats->kernel_ms = all_ms;
    stats->walk_ms = walk_ms;
    stats->transitions = sums[0];
    stats->work_items = static_cast<uint64_t>(src_end - src_begin) * n_walks;
    stats->e2e_ms = now_ms() - t0;
  }
  return TGTP_OK;
}

extern "C" int tgtp_mt_walks(tgtp_ctx* ctx, const tgtp_dgraph* dg, uint32_t src_begin,
                             uint32_t src_end, uint32_t depth, uint32_t n_walks, int32_t variant,
                             uint32_t* counters_out, uint64_t* visit_count, uint64_t* checksum,
                             tgtp_stats* stats) {
  if (!ctx) return TGTP_E_ARG;
  int rc = mt_walks_check(ctx, dg, src_begin, src_end, n_walks, variant);
  if (rc) return rc;
  const double t0 = now_ms();
  TGTP_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  uint32_t* d_counters =
      static_cast<uint32_t*>(ctx_buffer(ctx, kBufCounters, dg->n * sizeof(uint32_t)));
  auto* d_sums = static_cast<unsigned long long*>(ctx_buffer(ctx, kBufSums, 64));
  if (!d_counters || !d_sums) return set_error(ctx, TGTP_E_NOMEM, "counter buffers");
  return mt_walks_run(ctx, dg, src_begin, src_end, depth, n_walks, variant, counters_out,
                      visit_count, checksum, stats, d_counters, d_sums, t0, true);
}

extern "C" int tgtp_mt_walks_noalloc(tgtp_ctx* ctx, const tgtp_dgraph* dg, uint32_t src_begin,
                                     uint32_t src_end, uint32_t depth, uint32_t n_walks,
                                     int32_t variant, uint32_t* counters_out,
                                     uint64_t* visit_count, uint64_t* checksum,
                                     tgtp_stats* stats) {
  if (!ctx) return TGTP_E_ARG;
  int rc = mt_walks_check(ctx, dg, src_begin, src_end, n_walks, variant);
  if (rc) return rc;
  if (dg->n > kNoAllocMaxN)
    return set_error(ctx, TGTP_E_UNSUPPORTED,
                     "tgtp_mt_walks_noalloc: the graph has " + std::to_string(dg->n) +
                         " vertices; its scratch holds counters for at most " +
                         std::to_string(kNoAllocMaxN) + " (use tgtp_mt_walks)");
  if (!ctx->scratch)
    return set_error(ctx, TGTP_E_HIP, "tgtp_mt_walks_noalloc: device scratch address unavailable");
  const double t0 = now_ms();
  TGTP_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  // counters from the start of this device's scratch, the sums after the
  // largest counter array (an offset of 16 MiB: 8-byte aligned); both are
  // cleared by the walk's first kernel
  auto* d_counters = static_cast<uint32_t*>(ctx->scratch);
  auto* d_sums = reinterpret_cast<unsigned long long*>(d_counters + kNoAllocMaxN);
  return mt_walks_run(ctx, dg, src_begin, src_end, depth, n_walks, variant, counters_out,
                      visit_count, checksum, stats, d_counters, d_sums, t0, false);
}

extern "C" int tgtp_saba_walks(tgtp_ctx* ctx, const tgtp_dgraph* dg, uint32_t src_begin,
                               uint32_t src_end, uint32_t depth, uint32_t n_walks,
                               uint32_t* counters_out, uint64_t* visit_count,
                               uint64_t* checksum, tgtp_stats* stats) {
  if (!ctx) return TGTP_E_ARG;
  if (!dg) return set_error(ctx, TGTP_E_ARG, "null device graph");
  // n_walks == 0 walks nothing: counters and sums are zero, as the
  // reference's driver gives (rw_experiment_saba.cpp:160, 186)
  if (n_walks % 8 != 0)
    return set_error(ctx, TGTP_E_ARG,
                     "n_walks must be a multiple of 8 within the 32-bit grid domain");
  if (src_end > dg->n || src_begin > src_end)
    return set_error(ctx, TGTP_E_ARG, "source range outside [0, n]");
  const double t0 = now_ms();
  TGTP_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  uint32_t* d_counters =
      static_cast<uint32_t*>(ctx_buffer(ctx, kBufCounters, dg->n * sizeof(uint32_t)));
  auto* d_sums = static_cast<unsigned long long*>(ctx_buffer(ctx, kBufSums, 64));
  if (!d_counters || !d_sums) return set_error(ctx, TGTP_E_NOMEM, "counter buffers");
  float walk_ms = 0;
  int rc = launch_saba_walks(ctx, dg, src_begin, src_end, depth, n_walks, d_counters, d_sums,
                             &walk_ms);
  if (rc) return rc;
  unsigned long long sums[2] = {0, 0};
  TGTP_HIP_CHECK(ctx, hipMemcpyAsync(sums, d_sums, sizeof(sums), hipMemcpyDeviceToHost,
                                     ctx->stream));
  if (counters_out)
    TGTP_HIP_CHECK(ctx, hipMemcpyAsync(counters_out, d_counters, dg->n * sizeof(uint32_t),
                                       hipMemcpyDeviceToHost, ctx->stream));
  TGTP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (visit_count) *visit_count = sums[0];
  if (checksum) *checksum = sums[1];
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    float all_ms = 0;
    TGTP_HIP_CHECK(ctx, hipEventElapsedTime(&all_ms, ctx->ev[0], ctx->ev[3]));
    stats->kernel_ms = all_ms;
    stats->walk_ms = walk_ms;
    // transitions actually walked: sources with degree > 0 times W times depth
    stats->transitions = sums[0];
    stats->work_items = static_cast<uint64_t>(src_end - src_begin) * (n_walks / 8u);
    stats->e2e_ms = now_ms() - t0;
  }
  return TGTP_OK;
}
