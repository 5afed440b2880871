// tgtp_fast.hip -- FAST_FP32 / FAST_BF16 per-source pipeline.
//
// Same algorithm as AESC::tgtp (aesc.h:444-551) with the round structure
// made synchronous, so a whole 256-thread workgroup works on one source:
//   * diffusion round l -> l+1 (the synchronous form of aesc.h:219-239):
//     next frontier = neighbours of the frontier (bitmap, compacted in
//     vertex order), then pvec'[u] = (sum_{w in adj(u)} pvec[w]) / deg(u)
//     for every u in it.  pvec_l is then exactly P^l e_s (P = D^-1 A), the
//     series TGT+ truncates; the reference's in-place sweep pushes some mass
//     twice per round.  Each sum runs in adjacency order (one thread) or as a
//     fixed shuffle tree (one wave, high degree): deterministic.
//   * stopping rule, top-gamma bound, chi and walk counts as the reference
//     (RWRATIO 25, aesc.h:493-521); ties at the gamma-th value keep every
//     tied vertex (no permutation to reproduce).
//   * 2-way SABA walks with the reference's grid / hash / fallback per
//     bouquet; each bouquet's 8 lane sums are added in FP64, bouquets in a
//     fixed order; the per-edge row sums (preds) are FP64 throughout and
//     rounded to float once, after symmetrization.
//
// The CSR may be the disjoint union of a batch of graphs: the persistent
// workgroups drain one work list of sources across all of them (largest
// graphs first), each source reading its own graph's n, epsilon, gamma and
// log(m/delta), and keeping pvec / bitmap in LDS when its graph fits.
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "common.hpp"
#include "tgtp_fast.hpp"
#include "wave.hpp"

using namespace tgtp_dev;

namespace {

constexpr int kFB = 256;         // threads per source
constexpr int kFW = kFB / 64;    // waves per source
constexpr double kMaxWalksFast = 68719476736.0;
constexpr uint32_t kLdsCap = 64 * 1024;

struct FastShared {
  float rf[kFW];
  double rd[kFW];
  uint32_t ru[kFW];
  uint64_t r64[kFW];
  uint32_t heavy[kFB];
  uint32_t n_heavy;
};

// The per-source view: vertex ids are global (CSR), pvec / bitmap indices
// are local to the source's graph (global - off).
struct SrcView {
  uint32_t off, n;
  float* cur;
  float* nxt;
  uint32_t* bits;
};

__device__ __forceinline__ void set_bit(uint32_t* bits, uint32_t i) {
  atomicOr(&bits[i >> 5], 1u << (i & 31u));
}
__device__ __forceinline__ bool get_bit(const uint32_t* bits, uint32_t i) {
  return (bits[i >> 5] >> (i & 31u)) & 1u;
}

// Sets the bit of every neighbour of the frontier q[0, Q).
__device__ void mark_frontier(const FastArgs& A, const SrcView& V, const uint32_t* q, uint32_t Q,
                              FastShared& sh) {
  for (uint32_t base = 0; base < Q; base += kFB) {
    if (threadIdx.x == 0) sh.n_heavy = 0;
    __syncthreads();
    const uint32_t i = base + threadIdx.x;
    if (i < Q) {
      const uint32_t v = q[i];
      const uint2 rv = A.rowd[v];
      if (rv.y <= A.light) {
        for (uint32_t k = 0; k < rv.y; ++k) set_bit(V.bits, A.adj[rv.x + k] - V.off);
      } else {
        sh.heavy[atomicAdd(&sh.n_heavy, 1u)] = v;
      }
    }
    __syncthreads();
    const uint32_t nh = sh.n_heavy;
    for (uint32_t h = wave_id(); h < nh; h += kFW) {
      const uint2 rv = A.rowd[sh.heavy[h]];
      for (uint32_t k = lane_id(); k < rv.y; k += 64) set_bit(V.bits, A.adj[rv.x + k] - V.off);
    }
    __syncthreads();
  }
}

// Bitmap -> vertex-ordered list qn (global ids); clears the bitmap.
__device__ uint32_t compact_frontier(const SrcView& V, uint32_t* qn, FastShared& sh) {
  const uint32_t nwords = (V.n + 31) / 32;
  const uint32_t per = (nwords + kFB - 1) / kFB;
  const uint32_t w0 = min(nwords, threadIdx.x * per);
  const uint32_t w1 = min(nwords, w0 + per);
  uint32_t cnt = 0;
  for (uint32_t w = w0; w < w1; ++w) cnt += __popc(V.bits[w]);
  uint32_t total = 0;
  uint32_t off = block_exclusive_scan<kFW>(cnt, sh.ru, &total);
  for (uint32_t w = w0; w < w1; ++w) {
    uint32_t x = V.bits[w];
    while (x) {
      const uint32_t b = __ffs(x) - 1;
      x &= x - 1;
      qn[off++] = V.off + w * 32u + b;
    }
    V.bits[w] = 0u;
  }
  __syncthreads();
  return total;
}

// nxt[u] = sum(cur over adj(u)) / deg(u) for u in qn; per-thread partial
// max / min / degree sum through the references.
__device__ void pull_round(const FastArgs& A, const SrcView& V, const uint32_t* qn, uint32_t Qn,
                           FastShared& sh, float& mx, float& mn, uint64_t& dsum) {
  for (uint32_t base = 0; base < Qn; base += kFB) {
    if (threadIdx.x == 0) sh.n_heavy = 0;
    __syncthreads();
    const uint32_t i = base + threadIdx.x;
    if (i < Qn) {
      const uint32_t u = qn[i];
      const uint2 ru = A.rowd[u];
      dsum += ru.y;
      if (ru.y <= A.light) {
        float s = 0.0f;
        for (uint32_t k = 0; k < ru.y; ++k) s += V.cur[A.adj[ru.x + k] - V.off];
        const float val = s / A.degf[u];
        V.nxt[u - V.off] = val;
        mx = fmaxf(mx, val);
        mn = fminf(mn, val);
      } else {
        sh.heavy[atomicAdd(&sh.n_heavy, 1u)] = u;
      }
    }
    __syncthreads();
    const uint32_t nh = sh.n_heavy;
    for (uint32_t h = wave_id(); h < nh; h += kFW) {
      const uint32_t u = sh.heavy[h];
      const uint2 ru = A.rowd[u];
      float s = 0.0f;
      for (uint32_t k = lane_id(); k < ru.y; k += 64) s += V.cur[A.adj[ru.x + k] - V.off];
      s = wave_sum(s);
      if (lane_id() == 0) {
        const float val = s / A.degf[u];
        V.nxt[u - V.off] = val;
        mx = fmaxf(mx, val);
        mn = fminf(mn, val);
      }
    }
    __syncthreads();
  }
}

// One bouquet of the 2-way walk for lane b of an 8-lane group; returns the
// lane's float sum over the bouquet's hops.  The group's 8 lanes run in
// lockstep (same position, bouquet and hop counts).
__device__ __forceinline__ float walk2_bouquet(const FastArgs& A, const float* pv, uint32_t off,
                                               uint32_t s, uint32_t v, uint32_t j, uint32_t b,
                                               uint32_t step, uint32_t len, uint32_t giw) {
  uint32_t c1 = s, c2 = v;
  uint32_t s1 = (j * 8u + b + 1u) * step;
  uint32_t s2 = s1, st = s1;
  bool fb = false;
  float x = 0.0f;
  for (uint32_t h = 0; h < len; ++h) {
    const uint64_t z = __ballot(!fb && (s1 == 0u || s2 == 0u));
    fb = fb || (((z >> (giw * 8u)) & 0xFFull) != 0ull);
    const uint2 d1 = A.rowd[c1];
    const uint2 d2 = A.rowd[c2];
    uint32_t k1, k2;
    if (!fb) {
      k1 = __umulhi(s1 ^ fmix32(c1 - off), d1.y);
      k2 = __umulhi(s2 ^ fmix32(c2 - off), d2.y);
      s1 *= d1.y;
      s2 *= d2.y;
    } else {
      const uint32_t ra = splitmix32(st);
      const uint32_t rb = splitmix32(st);
      k1 = ra % d1.y;
      k2 = rb % d2.y;
    }
    c1 = A.adj[d1.x + k1];
    c2 = A.adj[d2.x + k2];
    x += pv[c1 - off] - pv[c2 - off];
  }
  return x;
}

// FP64 sum of the 8 lanes of a group, lanes 0..7 in order.
__device__ __forceinline__ double group_sum8(float x, uint32_t grp_base) {
  double s = 0.0;
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) s += static_cast<double>(__shfl(x, static_cast<int>(grp_base + k)));
  return s;
}

__device__ void fast_source(const FastArgs& A, uint32_t gi, uint32_t src, SrcView V, uint32_t* qa,
                            uint32_t* qb, FastShared& sh) {
  const uint32_t n = V.n, off = V.off;
  const double eps = A.g_eps[gi], logmd = A.g_logmd[gi];
  const int32_t gamma = A.g_gamma[gi];
  for (uint32_t i = threadIdx.x; i < n; i += kFB) {
    V.cur[i] = 0.0f;
    V.nxt[i] = 0.0f;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    V.cur[src - off] = 1.0f;
    qa[0] = src;
  }
  const uint2 rs = A.rowd[src];
  const uint32_t r0 = rs.x, dsi = rs.y, r1 = rs.x + rs.y;
  const double dsd = static_cast<double>(dsi);
  for (uint32_t p = r0 + threadIdx.x; p < r1; p += kFB) A.preds_d[p] = 1.0 / dsd;
  const double dse_d = dsd * eps;
  const double dse2 = dse_d * dse_d;
  uint32_t* q = qa;
  uint32_t* qn = qb;
  uint32_t Q = 1;
  int ell = 0;
  float gmax = 0.0f, gmin = 1.0f;
  __syncthreads();

  for (;;) {
    mark_frontier(A, V, q, Q, sh);
    const uint32_t Qn = compact_frontier(V, qn, sh);
    float mx = 0.0f, mn = 1.0f;
    uint64_t dsum = 0;
    pull_round(A, V, qn, Qn, sh, mx, mn, dsum);
    for (uint32_t i = threadIdx.x; i < Q; i += kFB) V.cur[q[i] - off] = 0.0f;
    __syncthreads();
    {
      float* t = V.cur;
      V.cur = V.nxt;
      V.nxt = t;
      uint32_t* tq = q;
      q = qn;
      qn = tq;
    }
    Q = Qn;
    ++ell;
    const double ps = static_cast<double>(V.cur[src - off]);
    for (uint32_t p = r0 + threadIdx.x; p < r1; p += kFB)
      A.preds_d[p] += (ps - static_cast<double>(V.cur[A.adj[p] - off])) / dsd;
    gmax = block_reduce<kFW>(mx, sh.rf, OpMaxF{});
    mn = block_reduce<kFW>(mn, sh.rf, OpMinF{});
    gmin = Q < n ? 0.0f : mn;
    const uint64_t dtot = block_reduce<kFW>(dsum, sh.r64, OpSum{});
    double rw = 0.0;
    for (uint32_t p = r0 + threadIdx.x; p < r1; p += kFB) {
      const int rest = A.taus[p] - ell;
      if (rest > 0) {
        const double chi = 2.0 * rest * (static_cast<double>(gmax) - gmin);
        rw += fmax(ceil(8 * (chi * chi) * logmd / dse2), 1.0);
      }
    }
    rw = block_reduce<kFW>(rw, sh.rd, OpSum{});
    if (static_cast<double>(dtot) >= 25 * rw) break;
  }

  // ---- top-gamma bound: gamma-th largest value by threshold passes
  const uint32_t rg = min(static_cast<uint32_t>(gamma), Q);
  float thr = INFINITY, vstar = 0.0f;
  uint32_t above = 0;
  for (uint32_t it = 0; it <= rg; ++it) {
    float m1 = -INFINITY;
    for (uint32_t i = threadIdx.x; i < Q; i += kFB) {
      const float x = V.cur[q[i] - off];
      if (x < thr) m1 = fmaxf(m1, x);
    }
    m1 = block_reduce<kFW>(m1, sh.rf, OpMaxF{});
    uint32_t c = 0;
    for (uint32_t i = threadIdx.x; i < Q; i += kFB) c += (V.cur[q[i] - off] == m1);
    c = block_reduce<kFW>(c, sh.ru, OpSum{});
    vstar = m1;
    if (above + c >= rg || c == 0) break;
    above += c;
    thr = m1;
  }
  const float gamma_max = gamma < static_cast<int>(Q) ? fminf(1.0f, vstar) : 0.0f;
  float emax = gmax + gamma_max;
  for (uint32_t i = threadIdx.x; i < Q; i += kFB) {
    const uint32_t u = q[i] - off;
    if (V.cur[u] >= vstar) set_bit(V.bits, u);
  }
  __syncthreads();
  {
    float best = -INFINITY;
    for (uint32_t base = 0; base < Q; base += kFB) {
      if (threadIdx.x == 0) sh.n_heavy = 0;
      __syncthreads();
      const uint32_t i = base + threadIdx.x;
      if (i < Q && V.cur[q[i] - off] >= vstar) sh.heavy[atomicAdd(&sh.n_heavy, 1u)] = q[i];
      __syncthreads();
      const uint32_t nh = sh.n_heavy;
      for (uint32_t h = wave_id(); h < nh; h += kFW) {
        const uint32_t u = sh.heavy[h];
        const uint2 ru = A.rowd[u];
        const float pu = V.cur[u - off];
        for (uint32_t k = lane_id(); k < ru.y; k += 64) {
          const uint32_t w = A.adj[ru.x + k] - off;
          if (get_bit(V.bits, w)) best = fmaxf(best, pu + V.cur[w]);
        }
      }
      __syncthreads();
    }
    best = block_reduce<kFW>(best, sh.rf, OpMaxF{});
    emax = fmaxf(emax, best);
  }
  for (uint32_t w = threadIdx.x; w < (n + 31) / 32; w += kFB) V.bits[w] = 0u;
  if (gamma == 1) emax = 2 * gmax;
  __syncthreads();

  // ---- chi and walk counts, one wave per CSR position (aesc.h:392-412, 508-521)
  float smx = 0.0f, smn = FLT_MAX;
  for (uint32_t p = r0 + threadIdx.x; p < r1; p += kFB) {
    const float x = V.cur[A.adj[p] - off];
    smx = fmaxf(smx, x);
    smn = fminf(smn, x);
  }
  smx = block_reduce<kFW>(smx, sh.rf, OpMaxF{});
  smn = block_reduce<kFW>(smn, sh.rf, OpMinF{});
  const float dse = static_cast<float>(dse_d);
  for (uint32_t p = r0 + wave_id(); p < r1; p += kFW) {
    const int rest = A.taus[p] - ell;
    if (rest <= 0) {
      if (lane_id() == 0) A.nr_pos[p] = 0;
      continue;
    }
    const uint2 rv = A.rowd[A.adj[p]];
    float tmx = 0.0f, tmn = FLT_MAX;
    for (uint32_t k = lane_id(); k < rv.y; k += 64) {
      const float x = V.cur[A.adj[rv.x + k] - off];
      tmx = fmaxf(tmx, x);
      tmn = fminf(tmn, x);
    }
    tmx = wave_max(tmx);
    tmn = wave_min(tmn);
    if (lane_id() == 0) {
      const int len = rest;
      const float chi = gmax + (smx + tmx) / 2.0 + (len - 1) * emax - smn - tmn -
                        2 * (len - 1) * gmin;
      const double v = fmax(ceil(8 * (chi * chi) * logmd / (dse * dse)), 1.0);
      uint64_t nr = 0;
      if (!(v < kMaxWalksFast)) {
        atomicOr(A.status, 1u);
      } else {
        nr = static_cast<uint64_t>(v);
        nr += 8 - nr % 8;
      }
      A.nr_pos[p] = nr;
    }
  }
  __syncthreads();

  // ---- 2-way walks.  Positions with at most small_bq bouquets: one 8-lane
  // group each (32 groups side by side); larger ones: the whole workgroup.
  const uint32_t g = threadIdx.x >> 3, b = threadIdx.x & 7u;
  const uint32_t giw = g & 7u, gbase = (threadIdx.x & 63u) & ~7u;
  const uint64_t small_bq = A.small_bq;
  for (uint32_t p = r0 + g; p < r1; p += kFB / 8) {
    const int rest = A.taus[p] - ell;
    const uint64_t nr = rest > 0 ? A.nr_pos[p] : 0;
    if (nr == 0 || nr / 8 > small_bq) continue;
    const uint32_t v = A.adj[p];
    const uint32_t step = static_cast<uint32_t>(0xFFFFFFFFull / nr);
    double acc = 0.0;
    for (uint32_t j = 0; j < nr / 8; ++j)
      acc += group_sum8(walk2_bouquet(A, V.cur, off, src, v, j, b, step, rest, giw), gbase);
    if (b == 0) {
      A.preds_d[p] += acc / (dsd * static_cast<double>(nr));
      if (A.tr_x) A.tr_x[p] = static_cast<float>(acc);
    }
  }
  __syncthreads();
  for (uint32_t p = r0; p < r1; ++p) {
    const int rest = A.taus[p] - ell;
    const uint64_t nr = rest > 0 ? A.nr_pos[p] : 0;
    if (nr / 8 <= small_bq) continue;
    const uint32_t v = A.adj[p];
    const uint32_t step = static_cast<uint32_t>(0xFFFFFFFFull / nr);
    const uint64_t nbq = nr / 8;
    double acc = 0.0;
    for (uint64_t j = g; j < nbq; j += kFB / 8)
      acc += group_sum8(walk2_bouquet(A, V.cur, off, src, v, static_cast<uint32_t>(j), b, step,
                                      rest, giw),
                        gbase);
    double mine = (b == 0) ? acc : 0.0;  // one contribution per group, groups in order
    mine = block_reduce<kFW>(mine, sh.rd, OpSum{});
    if (threadIdx.x == 0) {
      A.preds_d[p] += mine / (dsd * static_cast<double>(nr));
      if (A.tr_x) A.tr_x[p] = static_cast<float>(mine);
    }
  }
  if (threadIdx.x == 0) {
    if (A.tr_ell) A.tr_ell[src] = ell;
    if (A.tr_emax) A.tr_emax[src] = emax;
    if (A.tr_gmax) A.tr_gmax[src] = gmax;
    if (A.tr_gmin) A.tr_gmin[src] = gmin;
    if (A.tr_qsize) A.tr_qsize[src] = Q;
  }
  __syncthreads();
}

__device__ __forceinline__ uint32_t graph_of(const uint32_t* voff, uint32_t n_graphs, uint32_t v) {
  uint32_t lo = 0, hi = n_graphs;  // last k with voff[k] <= v
  while (hi - lo > 1) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (voff[mid] <= v)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(kFB) source_fast_kernel(FastArgs A) {
  extern __shared__ float lds_f[];
  __shared__ FastShared sh;
  __shared__ uint32_t src_sh;
  uint32_t* qa = A.qbuf + static_cast<uint64_t>(blockIdx.x) * 2ull * A.qcap;
  uint32_t* qb = qa + A.qcap;
  float* hbm = A.pv_g ? A.pv_g + static_cast<uint64_t>(blockIdx.x) * fast_slot_floats(A.slot_n)
                      : nullptr;
  // bitmaps start clear and every source leaves them clear
  for (uint32_t w = threadIdx.x; w < (A.lds_n + 31) / 32; w += kFB)
    reinterpret_cast<uint32_t*>(lds_f + 2 * A.lds_n)[w] = 0u;
  if (hbm)
    for (uint32_t w = threadIdx.x; w < (A.slot_n + 31) / 32; w += kFB)
      reinterpret_cast<uint32_t*>(hbm + 2 * A.slot_n)[w] = 0u;
  __syncthreads();
  for (;;) {
    if (threadIdx.x == 0) src_sh = atomicAdd(A.next_src, 1u);
    __syncthreads();
    const uint32_t k = src_sh;
    __syncthreads();
    if (k >= A.n_work) break;
    const uint32_t src = A.work ? A.work[k] : k;
    const uint32_t gi = graph_of(A.voff, A.n_graphs, src);
    SrcView V;
    V.off = A.voff[gi];
    V.n = A.voff[gi + 1] - V.off;
    if (V.n <= A.lds_n) {
      V.cur = lds_f;
      V.nxt = lds_f + A.lds_n;
      V.bits = reinterpret_cast<uint32_t*>(lds_f + 2 * A.lds_n);
    } else {
      V.cur = hbm;
      V.nxt = hbm + A.slot_n;
      V.bits = reinterpret_cast<uint32_t*>(hbm + 2 * A.slot_n);
    }
    fast_source(A, gi, src, V, qa, qb, sh);
  }
}

__global__ void __launch_bounds__(256)
symmetrize_d_kernel(const double* __restrict__ preds, const int32_t* __restrict__ src_of,
                    uint64_t m, float* __restrict__ out) {
  for (uint64_t t = blockIdx.x * 256ull + threadIdx.x; t < m; t += gridDim.x * 256ull) {
    const int32_t s = src_of[t];
    out[t] = static_cast<float>(preds[t] + (s >= 0 ? preds[s] : 0.0));
  }
}

}  // namespace

size_t fast_lds_bytes(uint64_t n) { return n ? fast_slot_floats(n) * 4 : 0; }

bool fast_fits_lds(uint64_t n) { return fast_lds_bytes(n) <= kLdsCap; }

int fast_slots_per_cu(uint32_t lds_n) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, reinterpret_cast<const void*>(&source_fast_kernel), kFB,
          fast_lds_bytes(lds_n)) != hipSuccess)
    per_cu = 1;
  return std::max(1, std::min(per_cu, 8));
}

int launch_fast_sources(tgtp_ctx* ctx, const FastArgs& A, uint32_t slots) {
  source_fast_kernel<<<slots, kFB, fast_lds_bytes(A.lds_n), ctx->stream>>>(A);
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  return TGTP_OK;
}

int launch_symmetrize_d(tgtp_ctx* ctx, const double* preds, const int32_t* src_of, uint64_t m,
                        float* out) {
  const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>(4096, (m + 255) / 256));
  symmetrize_d_kernel<<<static_cast<uint32_t>(blocks), 256, 0, ctx->stream>>>(preds, src_of, m, out);
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  return TGTP_OK;
}
