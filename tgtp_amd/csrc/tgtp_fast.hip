// tgtp_fast.hip -- FAST_FP32 / FAST_BF16 per-source pipeline.
//
// Same algorithm as AESC::tgtp (aesc.h:444-551) with the round structure
// made synchronous, so a whole 256-thread workgroup works on one source:
//   * diffusion round l -> l+1 (the synchronous form of aesc.h:219-239):
//     next frontier = neighbours of the frontier (LDS bitmap, compacted in
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
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "common.hpp"
#include "tgtp_fast.hpp"
#include "wave.hpp"

using namespace tgtp_dev;

namespace {

constexpr int kFB = 256;          // threads per source
constexpr int kFW = kFB / 64;     // waves per source
constexpr uint32_t kLight = 32;   // degree up to which one thread handles a vertex
constexpr double kMaxWalksFast = 68719476736.0;

struct FastShared {
  float rf[kFW];
  double rd[kFW];
  uint32_t ru[kFW];
  uint64_t r64[kFW];
  uint32_t heavy[kFB];
  uint32_t n_heavy;
};

// Sets bit w for every neighbour w of the frontier q[0, Q).
__device__ void mark_frontier(const FastArgs& A, const uint32_t* q, uint32_t Q, uint32_t* bits,
                              FastShared& sh) {
  for (uint32_t base = 0; base < Q; base += kFB) {
    if (threadIdx.x == 0) sh.n_heavy = 0;
    __syncthreads();
    const uint32_t i = base + threadIdx.x;
    if (i < Q) {
      const uint32_t v = q[i];
      const uint2 rv = A.rowd[v];
      if (rv.y <= kLight) {
        for (uint32_t k = 0; k < rv.y; ++k) {
          const uint32_t w = A.adj[rv.x + k];
          atomicOr(&bits[w >> 5], 1u << (w & 31u));
        }
      } else {
        sh.heavy[atomicAdd(&sh.n_heavy, 1u)] = v;
      }
    }
    __syncthreads();
    const uint32_t nh = sh.n_heavy;
    for (uint32_t h = wave_id(); h < nh; h += kFW) {
      const uint2 rv = A.rowd[sh.heavy[h]];
      for (uint32_t k = lane_id(); k < rv.y; k += 64) {
        const uint32_t w = A.adj[rv.x + k];
        atomicOr(&bits[w >> 5], 1u << (w & 31u));
      }
    }
    __syncthreads();
  }
}

// Bitmap -> vertex-ordered list qn; clears the bitmap.  Returns the count.
__device__ uint32_t compact_frontier(const uint32_t* bits_ro, uint32_t* bits, uint32_t nwords,
                                     uint32_t* qn, FastShared& sh) {
  const uint32_t per = (nwords + kFB - 1) / kFB;
  const uint32_t w0 = threadIdx.x * per;
  const uint32_t w1 = min(nwords, w0 + per);
  uint32_t cnt = 0;
  for (uint32_t w = w0; w < w1; ++w) cnt += __popc(bits_ro[w]);
  uint32_t total = 0;
  uint32_t off = block_exclusive_scan<kFW>(cnt, sh.ru, &total);
  for (uint32_t w = w0; w < w1; ++w) {
    uint32_t x = bits[w];
    while (x) {
      const uint32_t b = __ffs(x) - 1;
      x &= x - 1;
      qn[off++] = w * 32u + b;
    }
    bits[w] = 0u;
  }
  __syncthreads();
  return total;
}

// nxt[u] = sum(cur over adj(u)) / deg(u) for u in qn; returns per-thread
// partial max / min / degree sum through the references.
__device__ void pull_round(const FastArgs& A, const float* cur, float* nxt, const uint32_t* qn,
                           uint32_t Qn, FastShared& sh, float& mx, float& mn, uint64_t& dsum) {
  for (uint32_t base = 0; base < Qn; base += kFB) {
    if (threadIdx.x == 0) sh.n_heavy = 0;
    __syncthreads();
    const uint32_t i = base + threadIdx.x;
    if (i < Qn) {
      const uint32_t u = qn[i];
      const uint2 ru = A.rowd[u];
      dsum += ru.y;
      if (ru.y <= kLight) {
        float s = 0.0f;
        for (uint32_t k = 0; k < ru.y; ++k) s += cur[A.adj[ru.x + k]];
        const float val = s / A.degf[u];
        nxt[u] = val;
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
      for (uint32_t k = lane_id(); k < ru.y; k += 64) s += cur[A.adj[ru.x + k]];
      s = wave_sum(s);
      if (lane_id() == 0) {
        const float val = s / A.degf[u];
        nxt[u] = val;
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
__device__ __forceinline__ float walk2_bouquet(const FastArgs& A, const float* pv, uint32_t s,
                                               uint32_t v, uint32_t j, uint32_t b, uint32_t step,
                                               uint32_t len, uint32_t grp_in_wave) {
  uint32_t c1 = s, c2 = v;
  uint32_t s1 = (j * 8u + b + 1u) * step;
  uint32_t s2 = s1, st = s1;
  bool fb = false;
  float x = 0.0f;
  for (uint32_t h = 0; h < len; ++h) {
    const uint64_t z = __ballot(!fb && (s1 == 0u || s2 == 0u));
    fb = fb || (((z >> (grp_in_wave * 8u)) & 0xFFull) != 0ull);
    const uint2 d1 = A.rowd[c1];
    const uint2 d2 = A.rowd[c2];
    uint32_t k1, k2;
    if (!fb) {
      k1 = __umulhi(s1 ^ fmix32(c1), d1.y);
      k2 = __umulhi(s2 ^ fmix32(c2), d2.y);
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
    x += pv[c1] - pv[c2];
  }
  return x;
}

// FP64 sum of the 8 lanes of a group, lanes 0..7 in order (all lanes get it).
__device__ __forceinline__ double group_sum8(float x, uint32_t grp_base) {
  double s = 0.0;
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) s += static_cast<double>(__shfl(x, static_cast<int>(grp_base + k)));
  return s;
}

template <bool kLds>
__device__ void fast_source(const FastArgs& A, uint32_t src, float* cur, float* nxt,
                            uint32_t* bits, uint32_t* qa, uint32_t* qb, FastShared& sh) {
  const uint32_t n = A.n;
  const uint32_t nwords = (n + 31) / 32;
  for (uint32_t i = threadIdx.x; i < n; i += kFB) {
    cur[i] = 0.0f;
    nxt[i] = 0.0f;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    cur[src] = 1.0f;
    qa[0] = src;
  }
  const uint2 rs = A.rowd[src];
  const uint32_t r0 = rs.x, dsi = rs.y, r1 = rs.x + rs.y;
  const float ds = static_cast<float>(dsi);
  const double dsd = static_cast<double>(dsi);
  for (uint32_t p = r0 + threadIdx.x; p < r1; p += kFB) A.preds_d[p] = 1.0 / dsd;
  const double dse_d = dsd * A.eps;
  const double dse2 = dse_d * dse_d;
  uint32_t* q = qa;
  uint32_t* qn = qb;
  uint32_t Q = 1;
  int ell = 0;
  float gmax = 0.0f, gmin = 1.0f;
  __syncthreads();

  for (;;) {
    mark_frontier(A, q, Q, bits, sh);
    const uint32_t Qn = compact_frontier(bits, bits, nwords, qn, sh);
    float mx = 0.0f, mn = 1.0f;
    uint64_t dsum = 0;
    pull_round(A, cur, nxt, qn, Qn, sh, mx, mn, dsum);
    for (uint32_t i = threadIdx.x; i < Q; i += kFB) cur[q[i]] = 0.0f;
    __syncthreads();
    {
      float* t = cur;
      cur = nxt;
      nxt = t;
      uint32_t* tq = q;
      q = qn;
      qn = tq;
    }
    Q = Qn;
    ++ell;
    const double ps = static_cast<double>(cur[src]);
    for (uint32_t p = r0 + threadIdx.x; p < r1; p += kFB)
      A.preds_d[p] += (ps - static_cast<double>(cur[A.adj[p]])) / dsd;
    gmax = block_reduce<kFW>(mx, sh.rf, OpMaxF{});
    mn = block_reduce<kFW>(mn, sh.rf, OpMinF{});
    gmin = Q < n ? 0.0f : mn;
    const uint64_t dtot = block_reduce<kFW>(dsum, sh.r64, OpSum{});
    double rw = 0.0;
    for (uint32_t p = r0 + threadIdx.x; p < r1; p += kFB) {
      const int rest = A.taus[p] - ell;
      if (rest > 0) {
        const double chi = 2.0 * rest * (static_cast<double>(gmax) - gmin);
        rw += fmax(ceil(8 * (chi * chi) * A.logmd / dse2), 1.0);
      }
    }
    rw = block_reduce<kFW>(rw, sh.rd, OpSum{});
    if (static_cast<double>(dtot) >= 25 * rw) break;
  }

  // ---- top-gamma bound: gamma-th largest value by threshold passes
  const uint32_t rg = min(static_cast<uint32_t>(A.gamma), Q);
  float thr = INFINITY, vstar = 0.0f;
  uint32_t above = 0;
  for (uint32_t it = 0; it <= rg; ++it) {
    float m1 = -INFINITY;
    for (uint32_t i = threadIdx.x; i < Q; i += kFB) {
      const float x = cur[q[i]];
      if (x < thr) m1 = fmaxf(m1, x);
    }
    m1 = block_reduce<kFW>(m1, sh.rf, OpMaxF{});
    uint32_t c = 0;
    for (uint32_t i = threadIdx.x; i < Q; i += kFB) c += (cur[q[i]] == m1);
    c = block_reduce<kFW>(c, sh.ru, OpSum{});
    vstar = m1;
    if (above + c >= rg || c == 0) break;
    above += c;
    thr = m1;
  }
  float gamma_max = static_cast<int>(A.gamma) < static_cast<int>(Q) ? fminf(1.0f, vstar) : 0.0f;
  float emax = gmax + gamma_max;
  for (uint32_t i = threadIdx.x; i < Q; i += kFB) {
    const uint32_t u = q[i];
    if (cur[u] >= vstar) atomicOr(&bits[u >> 5], 1u << (u & 31u));
  }
  __syncthreads();
  {
    float best = -INFINITY;
    for (uint32_t base = 0; base < Q; base += kFB) {
      if (threadIdx.x == 0) sh.n_heavy = 0;
      __syncthreads();
      const uint32_t i = base + threadIdx.x;
      if (i < Q && cur[q[i]] >= vstar) sh.heavy[atomicAdd(&sh.n_heavy, 1u)] = q[i];
      __syncthreads();
      const uint32_t nh = sh.n_heavy;
      for (uint32_t h = wave_id(); h < nh; h += kFW) {
        const uint32_t u = sh.heavy[h];
        const uint2 ru = A.rowd[u];
        const float pu = cur[u];
        for (uint32_t k = lane_id(); k < ru.y; k += 64) {
          const uint32_t w = A.adj[ru.x + k];
          if (bits[w >> 5] & (1u << (w & 31u))) best = fmaxf(best, pu + cur[w]);
        }
      }
      __syncthreads();
    }
    best = block_reduce<kFW>(best, sh.rf, OpMaxF{});
    emax = fmaxf(emax, best);
  }
  for (uint32_t w = threadIdx.x; w < nwords; w += kFB) bits[w] = 0u;
  if (A.gamma == 1) emax = 2 * gmax;
  __syncthreads();

  // ---- chi and walk counts, one wave per CSR position (aesc.h:392-412, 508-521)
  float smx = 0.0f, smn = FLT_MAX;
  for (uint32_t p = r0 + threadIdx.x; p < r1; p += kFB) {
    const float x = cur[A.adj[p]];
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
      const float x = cur[A.adj[rv.x + k]];
      tmx = fmaxf(tmx, x);
      tmn = fminf(tmn, x);
    }
    tmx = wave_max(tmx);
    tmn = wave_min(tmn);
    if (lane_id() == 0) {
      const int len = rest;
      const float chi = gmax + (smx + tmx) / 2.0 + (len - 1) * emax - smn - tmn -
                        2 * (len - 1) * gmin;
      const double v = fmax(ceil(8 * (chi * chi) * A.logmd / (dse * dse)), 1.0);
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

  // ---- 2-way walks.  Positions with at most kSmallBq bouquets: one 8-lane
  // group each (32 groups side by side); larger ones: the whole workgroup.
  constexpr uint64_t kSmallBq = 8;
  const uint32_t g = threadIdx.x >> 3, b = threadIdx.x & 7u;
  const uint32_t giw = g & 7u, gbase = (threadIdx.x & 63u) & ~7u;
  for (uint32_t p = r0 + g; p < r1; p += kFB / 8) {
    const int rest = A.taus[p] - ell;
    const uint64_t nr = rest > 0 ? A.nr_pos[p] : 0;
    if (nr == 0 || nr / 8 > kSmallBq) continue;
    const uint32_t v = A.adj[p];
    const uint32_t step = static_cast<uint32_t>(0xFFFFFFFFull / nr);
    double acc = 0.0;
    for (uint32_t j = 0; j < nr / 8; ++j)
      acc += group_sum8(walk2_bouquet(A, cur, src, v, j, b, step, rest, giw), gbase);
    if (b == 0) {
      A.preds_d[p] += acc / (dsd * static_cast<double>(nr));
      if (A.tr_x) A.tr_x[p] = static_cast<float>(acc);
    }
  }
  __syncthreads();
  for (uint32_t p = r0; p < r1; ++p) {
    const int rest = A.taus[p] - ell;
    const uint64_t nr = rest > 0 ? A.nr_pos[p] : 0;
    if (nr / 8 <= kSmallBq) continue;
    const uint32_t v = A.adj[p];
    const uint32_t step = static_cast<uint32_t>(0xFFFFFFFFull / nr);
    const uint64_t nbq = nr / 8;
    double acc = 0.0;
    for (uint64_t j = g; j < nbq; j += kFB / 8)
      acc += group_sum8(walk2_bouquet(A, cur, src, v, static_cast<uint32_t>(j), b, step, rest, giw),
                        gbase);
    // groups in order: lane 0 of each group contributes once
    double mine = (b == 0) ? acc : 0.0;
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

template <bool kLds>
__global__ void __launch_bounds__(kFB) source_fast_kernel(FastArgs A) {
  extern __shared__ float lds_f[];
  __shared__ FastShared sh;
  const uint32_t n = A.n;
  float *cur, *nxt;
  uint32_t* bits;
  if constexpr (kLds) {
    cur = lds_f;
    nxt = lds_f + n;
    bits = reinterpret_cast<uint32_t*>(lds_f + 2 * n);
  } else {
    const uint64_t per = 2ull * n + (n + 31) / 32;
    cur = A.pv_g + static_cast<uint64_t>(blockIdx.x) * per;
    nxt = cur + n;
    bits = reinterpret_cast<uint32_t*>(nxt + n);
  }
  uint32_t* qa = A.qbuf + static_cast<uint64_t>(blockIdx.x) * 2ull * n;
  uint32_t* qb = qa + n;
  for (uint32_t w = threadIdx.x; w < (n + 31) / 32; w += kFB) bits[w] = 0u;
  __syncthreads();
  __shared__ uint32_t src_sh;
  for (;;) {
    if (threadIdx.x == 0) src_sh = atomicAdd(A.next_src, 1u);
    __syncthreads();
    const uint32_t src = src_sh;
    __syncthreads();
    if (src >= n) break;
    fast_source<kLds>(A, src, cur, nxt, bits, qa, qb, sh);
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

size_t fast_lds_bytes(uint64_t n) { return n * 8 + ((n + 31) / 32) * 4; }

bool fast_fits_lds(uint64_t n) { return fast_lds_bytes(n) <= 64 * 1024; }

int launch_fast_sources(tgtp_ctx* ctx, const FastArgs& A, uint32_t slots, bool use_lds) {
  hipStream_t s = ctx->stream;
  if (use_lds)
    source_fast_kernel<true><<<slots, kFB, fast_lds_bytes(A.n), s>>>(A);
  else
    source_fast_kernel<false><<<slots, kFB, 0, s>>>(A);
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

int fast_slots_per_cu(bool use_lds, uint64_t n) {
  int per_cu = 0;
  const void* fn = use_lds ? reinterpret_cast<const void*>(&source_fast_kernel<true>)
                           : reinterpret_cast<const void*>(&source_fast_kernel<false>);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kFB,
                                                   use_lds ? fast_lds_bytes(n) : 0) != hipSuccess)
    per_cu = 1;
  return std::max(1, std::min(per_cu, use_lds ? 8 : 4));
}
