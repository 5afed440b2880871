// tgtp.hip -- TGT+ on gfx950: tau bounds, per-source push diffusion with the
// walk-cost stopping rule, top-gamma edge bound, chi / walk counts, 2-way
// SABA walks and symmetrization (AESC::tgtp, aesc.h:120-127, 164-216,
// 218-551).
//
// FAITHFUL mode evaluates every expression with the reference's own mix of
// float and double and in the reference's order, so the scores equal an
// IEEE build of the reference bit for bit:
//   * tau: one thread per undirected edge runs the serial refinement loop;
//     the only transcendental on the device is log() inside ceil(a/b - 1),
//     and any edge whose a/b - 1 lands within 1e-10 (relative) of an integer
//     is recomputed on the host with the C library's log (host_tgtp.cpp).
//   * per source: one 64-lane wave owns the source.  The in-place push
//     (aesc.h:222-234) runs over the frontier in order, each frontier
//     vertex's adjacency spread across the lanes: targets within one row are
//     distinct except for duplicate edges, which are adjacent and add the
//     same amount, so each duplicate run is folded by its first lane and the
//     result equals the sequential loop.  First-discovery order of the next
//     frontier is kept with a ballot prefix.  pvec lives in LDS when it fits.
//   * top-gamma: threshold selection by wave reductions; when values tie
//     across the gamma-th slot the exact std::nth_element permutation is
//     reproduced (select.hpp) on one lane.
//   * 2-way walks: 8 lanes per CSR position (eight positions per wave); each
//     lane accumulates its bouquet lane over all bouquets in order, exactly
//     as the reference's per-lane AVX2 accumulator, then lanes 0..7 are
//     summed in order.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "host_tgtp.hpp"
#include "select.hpp"
#include "tgtp_internal.hpp"
#include "taus_fast.hpp"
#include "tgtp_fast.hpp"

namespace {

constexpr double kNearInt = 1e-10;
// walks per CSR position above which a call is refused (2^36 walks of one
// edge would run for hours on any device)
constexpr double kMaxWalks = 68719476736.0;

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }
__device__ __forceinline__ uint64_t lanes_below() { return (1ull << lane_id()) - 1ull; }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
  return v;
}
__device__ __forceinline__ float wave_min(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
  return v;
}
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t o = __shfl_xor(v, off);
    v = o > v ? o : v;
  }
  return v;
}

// ---- tau bounds ---------------------------------------------------------------

struct TauArgs {
  const uint32_t* idx;
  const uint32_t* adj;
  const float* X;       // [n*omega]
  const double* onep;   // 1 + lambda_i
  const double* onem;   // 1 - lambda_i
  const double* pw;     // [k_rows][omega] lambda_i^(2(k+1))
  uint64_t m;
  uint32_t n;
  int32_t omega;
  uint32_t k_rows;
  float eps;            // float(epsilon / 2)
  double sq1m_1, b_1;   // 1 - lambda_1^2, log(1/|lambda_1|)
  double sq1m_w, b_w;   // same for lambda_{omega-1}
  int32_t* taus;
  uint32_t* flags;      // positions needing the host
  uint32_t* n_flags;
  uint32_t* tau0_max;
};

// get_tau (aesc.h:129-135); `near` is raised when ceil() is within rounding
// distance of flipping, the only place a device log() may differ from glibc.
__device__ int tau_bound_dev(double du, double dv, float eps, double sq1m, double b, float del,
                             float ups, bool& near) {
  const float eps_del = eps - del > 0.0 ? eps - del : eps;
  const double x = (1.0 / du + 1.0 / dv - 2.0 / du / dv - ups) / eps_del / sq1m;
  const double xm = (x < 1.0) ? 1.0 : x;  // std::max(x, 1.0)
  const double a = log(xm);
  const double y = a / b - 1;
  if (xm != 1.0 && fabs(y - rint(y)) <= kNearInt * fmax(1.0, fabs(y))) near = true;
  const double c = ceil(y);
  const int tau = static_cast<int>((c < 1.0) ? 1.0 : c);
  return tau % 2 == 0 ? tau + 1 : tau;
}

__device__ __forceinline__ uint32_t row_of(const uint32_t* idx, uint32_t n, uint64_t p) {
  // last u with idx[u] <= p (rows may be empty)
  uint32_t lo = 0, hi = n;  // answer in [lo, hi)
  while (hi - lo > 1) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (idx[mid] <= p)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint32_t first_pos(const uint32_t* idx, const uint32_t* adj,
                                              uint32_t v, uint32_t u) {
  uint32_t lo = idx[v], hi = idx[v + 1];
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (adj[mid] < u)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(256) tau0_kernel(TauArgs A) {
  uint32_t best = 0;
  for (uint64_t p = blockIdx.x * 256ull + threadIdx.x; p < A.m; p += gridDim.x * 256ull) {
    const uint32_t u = row_of(A.idx, A.n, p);
    const uint32_t v = A.adj[p];
    if (v < u) continue;
    bool near = false;
    const double du = A.idx[v + 1] - A.idx[v], dv = A.idx[u + 1] - A.idx[u];
    const int t0 = tau_bound_dev(du, dv, A.eps, A.sq1m_1, A.b_1, 0.0f, 0.0f, near);
    best = max(best, static_cast<uint32_t>(t0) + (near ? 2u : 0u));
  }
  for (int off = 32; off > 0; off >>= 1) best = max(best, (uint32_t)__shfl_xor(best, off));
  if ((threadIdx.x & 63u) == 0) atomicMax(A.tau0_max, best);
}

__global__ void __launch_bounds__(256) taus_faithful_kernel(TauArgs A) {
  for (uint64_t p = blockIdx.x * 256ull + threadIdx.x; p < A.m; p += gridDim.x * 256ull) {
    const uint32_t u = row_of(A.idx, A.n, p);
    const uint32_t v = A.adj[p];
    if (v < u) continue;
    const int om = A.omega;
    const float* xu = A.X + static_cast<uint64_t>(u) * om;
    const float* xv = A.X + static_cast<uint64_t>(v) * om;
    float ups = 0.0f;
    for (int i = 1; i < om - 1; ++i) {
      const float d = xu[i] - xv[i];
      ups = static_cast<float>(static_cast<double>(ups) + static_cast<double>(d * d) * A.onep[i]);
    }
    ups = ups / static_cast<float>(A.m);
    const double du = A.idx[v + 1] - A.idx[v], dv = A.idx[u + 1] - A.idx[u];
    bool near = false;
    int tau = tau_bound_dev(du, dv, A.eps, A.sq1m_1, A.b_1, 0.0f, 0.0f, near);
    int t = 1;
    uint32_t k = 0;
    for (;;) {
      if (k >= A.k_rows) {  // table too short: let the host finish this edge
        near = true;
        break;
      }
      const double* pw = A.pw + static_cast<uint64_t>(k) * om;
      float del = 0.0f;
      for (int i = 1; i < om - 1; ++i) {
        const float d = xu[i] - xv[i];
        const float pf = static_cast<float>(pw[i]);
        del = static_cast<float>(static_cast<double>(del) +
                                 static_cast<double>((d * d) * pf) / A.onem[i]);
      }
      del = del / static_cast<float>(A.m);
      const int tp = tau_bound_dev(du, dv, A.eps, A.sq1m_w, A.b_w, del, ups, near);
      if (t <= tp && tp < tau) {
        tau = tp;
        t += 2;
        ++k;
      } else {
        break;
      }
    }
    A.taus[p] = tau;
    A.taus[first_pos(A.idx, A.adj, v, u)] = tau;
    if (near) {
      const uint32_t slot = atomicAdd(A.n_flags, 1u);
      A.flags[slot] = static_cast<uint32_t>(p);
    }
  }
}

// ---- per-source pipeline --------------------------------------------------------

struct SrcArgs {
  const uint2* rowd;
  const uint32_t* adj;
  const float* degf;
  const int32_t* taus;
  float* preds;
  uint64_t* nr_pos;
  uint32_t* qbuf;     // [slots][2n + 128]: two queues of n entries, 64 spare words after each
  float* pvec_g;      // [slots][n]   (global-memory variant)
  uint8_t* seen_g;    // [slots][n]   (global-memory variant)
  uint32_t* next_src;
  uint32_t* n_ambiguous;
  uint32_t* status;   // bit 0: a walk count exceeded kMaxWalks
  const uint32_t* src_list;  // sources to run (null: all n, in id order)
  uint32_t n_src;            // entries of src_list
  // 2-way walks: 0 SABA (aesc.h:241-391); with the MT19937 comparator
  // (aesc_baseline.h:363-379) 1 = counting pass (walk counts only, no
  // walks), 2 = walk pass reading each position's draws from mt_draws at
  // mt_off[p]
  int32_t walk_mode;
  const uint32_t* mt_draws;
  const uint64_t* mt_off;
  uint32_t n;
  int32_t gamma;
  double eps;
  double logmd;       // log(m / delta), host glibc
  int32_t* tr_ell;
  float* tr_emax;
  float* tr_gmax;
  float* tr_gmin;
  uint32_t* tr_qsize;
  float* tr_x;
};

// Entries between a pipeline stage of the push's frontier loop and the stage
// it depends on (a power of two).
#ifndef TGTP_PIPE_GAP
#define TGTP_PIPE_GAP 2
#endif
constexpr uint32_t kPipeGap = TGTP_PIPE_GAP;
static_assert((kPipeGap & (kPipeGap - 1u)) == 0u, "a power of two");

template <bool kLds>
__device__ void faithful_source(const SrcArgs& A, uint32_t src, float* pvec, uint8_t* seen,
                                uint32_t* qa, uint32_t* qb) {
  const uint32_t ln = lane_id();
  const uint32_t n = A.n;
  for (uint32_t i = ln; i < n; i += 64) pvec[i] = 0.0f;
  __syncthreads();
  if (ln == 0) pvec[src] = 1.0f;  // aesc.h:461
  const uint2 rs = A.rowd[src];
  const uint32_t r0 = rs.x, dsi = rs.y, r1 = rs.x + rs.y;
  const float ds = static_cast<float>(dsi);
  const float init_pred = static_cast<float>(1.0 / static_cast<double>(ds));  // aesc.h:469
  for (uint32_t p = r0 + ln; p < r1; p += 64) A.preds[p] = init_pred;
  uint32_t* q = qa;
  uint32_t* qn = qb;
  if (ln == 0) q[0] = src;
  uint32_t qsize = 1;
  int ell = 0;
  float gmin = 1.0f, gmax = 0.0f;
  const double dse_d = static_cast<double>(ds) * A.eps;
  const double dse2 = dse_d * dse_d;  // pow(ds * eps, 2), aesc.h:498
  __syncthreads();

  // Ordering of the push's state between frontier vertices.  pvec / seen in
  // LDS: the workgroup is one wave and the LDS operations of a wave are
  // performed in order, so a compiler-level wave barrier keeps the program
  // order and no wait is needed (the prefetch loads below stay in flight).
  // pvec / seen in global memory: a full barrier, so that each store has
  // completed before a later load of the same word.
  auto vsync = [&]() {
    if constexpr (kLds) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
      __syncthreads();
    }
  };

  for (;;) {
    // ---- diffuse: aesc.h:222-234, frontier in order
    uint32_t qn_size = 0;
    // The queue, row descriptors, adjacency and receiver degrees do not
    // depend on pvec: a four-stage pipeline loads, for frontier entries ahead
    // of the current one, the vertex id, its row descriptor, this lane's
    // neighbour in the row's first 64 entries and that neighbour's degree,
    // each stage kPipeGap entries after the one it depends on, so a step waits
    // only for loads issued kPipeGap steps earlier.  Entry e's stages live in
    // slot e % kSlots and the loop is unrolled kSlots times, so that no
    // register holding a load in flight is copied, selected or moved to a
    // scalar register before its step (each of those waits for the load).
    // The loads are branch-free and their words are used as they come: an
    // entry past the frontier reads q[0] (the frontier is never empty: every
    // vertex has a neighbour), a lane past its row reads adj[0] and degf[0],
    // values the push ignores.  The queue index carries a zero the compiler
    // cannot see as uniform, which keeps the ids and row descriptors in
    // vector registers.
    uint32_t zv;
    __asm__ volatile("v_mov_b32 %0, 0" : "=v"(zv));
    auto qv = [&](uint32_t k) { return q[(k < qsize ? k : 0u) + zv]; };
    auto first_nb = [&](uint2 r) { return A.adj[ln < r.y ? r.x + ln : 0u]; };
    auto nb_deg = [&](uint2 r, uint32_t u) { return A.degf[ln < r.y ? u : 0u]; };
    // one chunk of 64 row entries of frontier vertex v: u is this lane's
    // entry (0xFFFFFFFF past the row), g its degree as a float
    // `before` is the row entry preceding the chunk (base > 0)
    auto push_chunk = [&](const uint2 rv, uint32_t base, uint32_t u, float residue, float g,
                          uint32_t before) {
      const uint32_t j = base + ln;
      const bool valid = j < rv.y;
      const uint32_t up = __shfl_up(u, 1);
      const bool lead = valid && (ln == 0 || u != up);
      bool newrun = lead;
      if (ln == 0 && base > 0 && valid) newrun = (u != before);
      const uint64_t leads = __ballot(lead);
      const uint32_t nvalid = min(64u, rv.y - base);
      uint32_t run = 0;
      if (lead) {
        const uint64_t higher = (ln == 63) ? 0ull : (leads >> (ln + 1));
        run = higher ? static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(higher)))
                     : (nvalid - ln);
      }
      const bool fresh = newrun && !seen[u];
      const uint64_t fm = __ballot(fresh);
      // every lane stores (the others to the queue's spare words past n),
      // so the count of memory operations does not depend on the data
      qn[fresh ? qn_size + __popcll(fm & lanes_below()) : n + ln] = u;
      if (fresh) seen[u] = 1;
      qn_size += __popcll(fm);
      if (lead) {
        const float c = residue / g;
        float t = pvec[u];
        for (uint32_t k = 0; k < run; ++k) t = t + c;
        pvec[u] = t;
      }
      vsync();
    };
    constexpr uint32_t kSlots = 4u * kPipeGap;
    uint32_t Vs[kSlots];
    uint2 Rs[kSlots];
    uint32_t Us[kSlots];
    float Gs[kSlots];
#pragma unroll
    for (uint32_t e = 0; e < kSlots; ++e) {
      Vs[e] = qv(e);
      Rs[e] = make_uint2(0u, 0u);
      Us[e] = 0u;
      Gs[e] = 1.0f;
    }
#pragma unroll
    for (uint32_t e = 0; e < 3u * kPipeGap; ++e) Rs[e] = A.rowd[min(Vs[e], n - 1u)];
#pragma unroll
    for (uint32_t e = 0; e < 2u * kPipeGap; ++e) Us[e] = first_nb(Rs[e]);
#pragma unroll
    for (uint32_t e = 0; e < kPipeGap; ++e) Gs[e] = nb_deg(Rs[e], Us[e]);
    for (uint32_t i0 = 0; i0 < qsize; i0 += kSlots) {
#pragma unroll
      for (uint32_t e = 0; e < kSlots; ++e) {
        const uint32_t i = i0 + e;
        if (i < qsize) {  // uniform
          const uint32_t v = __builtin_amdgcn_readfirstlane(Vs[e]);
          const uint2 rv = make_uint2(__builtin_amdgcn_readfirstlane(Rs[e].x),
                                      __builtin_amdgcn_readfirstlane(Rs[e].y));
          const uint32_t u_first = ln < rv.y ? Us[e] : 0xFFFFFFFFu;
          const float g_first = Gs[e];
          // stages of entries i + kPipeGap .. i + 4 kPipeGap
          constexpr uint32_t kM = kSlots - 1u;
          Gs[(e + kPipeGap) & kM] = nb_deg(Rs[(e + kPipeGap) & kM], Us[(e + kPipeGap) & kM]);
          Us[(e + 2u * kPipeGap) & kM] = first_nb(Rs[(e + 2u * kPipeGap) & kM]);
          Rs[(e + 3u * kPipeGap) & kM] = A.rowd[min(Vs[(e + 3u * kPipeGap) & kM], n - 1u)];
          Vs[e] = qv(i + kSlots);
          const float residue = pvec[v];
          vsync();
          if (ln == 0) pvec[v] = 0.0f;
          vsync();
          if (rv.y != 0) push_chunk(rv, 0u, u_first, residue, g_first, 0u);
          if (rv.y > 64u) {  // uniform
            // later chunks: each chunk's ids load during the chunk before
            // (branch-free: entries past the row read entry 0, dropped)
            auto ids_at = [&](uint32_t base) {
              return A.adj[base + ln < rv.y ? rv.x + base + ln : 0u];
            };
            uint32_t before = __shfl(u_first, 63);  // the first chunk is full
            uint32_t uc = ids_at(64u);
            for (uint32_t base = 64; base < rv.y; base += 64) {
              const bool in = base + ln < rv.y;
              const float g = A.degf[in ? uc : 0u];
              const uint32_t un = ids_at(base + 64u);
              const uint32_t u = in ? uc : 0xFFFFFFFFu;
              push_chunk(rv, base, u, residue, g, before);
              before = __shfl(uc, 63);
              uc = un;
            }
          }
        }
      }
    }
    // the next queue (global memory) is read back below: its stores complete
    __syncthreads();
    // preds update (aesc.h:235-237)
    const float ps = pvec[src];
    for (uint32_t p = r0 + ln; p < r1; p += 64) A.preds[p] += (ps - pvec[A.adj[p]]) / ds;
    for (uint32_t i = ln; i < qn_size; i += 64) seen[qn[i]] = 0;
    {
      uint32_t* t = q;
      q = qn;
      qn = t;
    }
    qsize = qn_size;
    ++ell;
    __syncthreads();
    // frontier statistics (aesc.h:484-491)
    uint64_t dsum = 0;
    float mx = 0.0f, mn = 1.0f;
    for (uint32_t i = ln; i < qsize; i += 64) {
      const uint32_t v = q[i];
      dsum += A.rowd[v].y;
      const float x = pvec[v];
      mx = fmaxf(mx, x);
      mn = fminf(mn, x);
    }
    dsum = wave_sum(dsum);
    gmax = wave_max(mx);
    mn = wave_min(mn);
    gmin = qsize < n ? 0.0f : mn;
    // walk-cost estimate (aesc.h:493-500): sums of integer-valued doubles,
    // exact in any order below 2^53
    double rw = 0.0;
    bool big = false;
    for (uint32_t p = r0 + ln; p < r1; p += 64) {
      const int rest = A.taus[p] - ell;
      if (rest > 0) {
        const double chi = 2.0 * rest * (static_cast<double>(gmax) - gmin);
        const double nr = fmax(ceil(8 * (chi * chi) * A.logmd / dse2), 1.0);
        big = big || nr >= 4503599627370496.0;
        rw += nr;
      }
    }
    rw = wave_sum(rw);
    if (__any(big)) {  // order matters above 2^53: repeat serially
      double s = 0.0;
      if (ln == 0) {
        for (uint32_t p = r0; p < r1; ++p) {
          const int rest = A.taus[p] - ell;
          if (rest <= 0) continue;
          const double chi = 2.0 * rest * (static_cast<double>(gmax) - gmin);
          s += fmax(ceil(8 * (chi * chi) * A.logmd / dse2), 1.0);
        }
      }
      rw = __shfl(s, 0);
    }
    const double picost = static_cast<double>(dsum);
    if (picost >= 25 * rw) break;  // RWRATIO
  }

  // ---- top-gamma edge bound (aesc.h:414-442, 504-506)
  const uint32_t rg = min(static_cast<uint32_t>(A.gamma), qsize);
  float thr = INFINITY, vstar = 0.0f;
  uint32_t above = 0, ties = 0;
  for (uint32_t it = 0; it <= rg; ++it) {  // each pass consumes >= 1 value
    float mx = -INFINITY;
    for (uint32_t i = ln; i < qsize; i += 64) {
      const float x = pvec[q[i]];
      if (x < thr) mx = fmaxf(mx, x);
    }
    mx = wave_max(mx);
    uint32_t cnt = 0;
    for (uint32_t i = ln; i < qsize; i += 64) cnt += (pvec[q[i]] == mx);
    cnt = wave_sum(cnt);
    if (above + cnt >= rg || cnt == 0) {  // cnt == 0 only for non-finite values
      vstar = mx;
      ties = cnt;
      break;
    }
    above += cnt;
    thr = mx;
  }
  // ties across the cut go to the exact permutation; values the threshold
  // scan cannot order (non-finite) have no defined order at all: the unguarded
  // partition must never see them, so the call reports an error instead
  const bool unordered = ties == 0;
  if (unordered && ln == 0) atomicOr(A.status, 2u);
  const bool ambiguous = !unordered && above + ties != rg;
  __syncthreads();
  if (ambiguous) {
    if (ln == 0) {
      tgtp_select::nth_element<uint32_t>(q, static_cast<int64_t>(rg) - 1, qsize,
                                         tgtp_select::Greater<uint32_t>{pvec});
      atomicAdd(A.n_ambiguous, 1u);
    }
    __syncthreads();
    for (uint32_t i = ln; i < rg; i += 64) seen[q[i]] = 1;
  } else {
    for (uint32_t i = ln; i < qsize; i += 64)
      if (pvec[q[i]] >= vstar) seen[q[i]] = 1;
  }
  __syncthreads();
  float gamma_max = fminf(1.0f, vstar);
  gamma_max = static_cast<int>(A.gamma) < static_cast<int>(qsize) ? gamma_max : 0.0f;
  float emax = gmax + gamma_max;
  {
    float best = -INFINITY;
    const uint32_t lim = ambiguous ? rg : qsize;
    for (uint32_t base = 0; base < lim; base += 64) {
      const uint32_t i = base + ln;
      bool member = false;
      uint32_t u = 0;
      if (i < lim) {
        u = q[i];
        member = ambiguous ? true : (pvec[u] >= vstar);
      }
      uint64_t mm = __ballot(member);
      while (mm) {
        const int l = __ffsll(static_cast<unsigned long long>(mm)) - 1;
        mm &= mm - 1;
        const uint32_t w0 = __shfl(u, l);
        const uint2 rw0 = A.rowd[w0];
        const float pu = pvec[w0];
        for (uint32_t k = ln; k < rw0.y; k += 64) {
          const uint32_t w = A.adj[rw0.x + k];
          if (seen[w]) best = fmaxf(best, pu + pvec[w]);
        }
      }
    }
    best = wave_max(best);
    emax = fmaxf(emax, best);
  }
  __syncthreads();
  if (ambiguous) {
    for (uint32_t i = ln; i < rg; i += 64) seen[q[i]] = 0;
  } else {
    for (uint32_t i = ln; i < qsize; i += 64)
      if (pvec[q[i]] >= vstar) seen[q[i]] = 0;
  }
  if (A.gamma == 1) emax = 2 * gmax;
  __syncthreads();

  // ---- chi and walk counts per CSR position (aesc.h:392-412, 508-521)
  float smx = 0.0f, smn = FLT_MAX;
  for (uint32_t p = r0 + ln; p < r1; p += 64) {
    const float x = pvec[A.adj[p]];
    smx = fmaxf(smx, x);
    smn = fminf(smn, x);
  }
  smx = wave_max(smx);
  smn = wave_min(smn);
  const float dse = static_cast<float>(static_cast<double>(ds) * A.eps);
  for (uint32_t p = r0; p < r1; ++p) {
    const int rest = A.taus[p] - ell;
    if (rest <= 0) {
      if (ln == 0) A.nr_pos[p] = 0;
      continue;
    }
    const uint2 rv = A.rowd[A.adj[p]];
    float tmx = 0.0f, tmn = FLT_MAX;
    for (uint32_t k = ln; k < rv.y; k += 64) {
      const float x = pvec[A.adj[rv.x + k]];
      tmx = fmaxf(tmx, x);
      tmn = fminf(tmn, x);
    }
    tmx = wave_max(tmx);
    tmn = wave_min(tmn);
    if (ln == 0) {
      const int len = rest;
      const float chi = gmax + (smx + tmx) / 2.0 + (len - 1) * emax - smn - tmn -
                        2 * (len - 1) * gmin;
      const double v = fmax(ceil(8 * (chi * chi) * A.logmd / (dse * dse)), 1.0);
      uint64_t nr;
      if (!(v < kMaxWalks)) {  // beyond any finishing run (or not finite): refuse
        atomicOr(A.status, 1u);
        nr = 0;
      } else {
        nr = static_cast<uint64_t>(v);
        nr += 8 - nr % 8;
      }
      A.nr_pos[p] = nr;
    }
  }
  __syncthreads();

  if (A.walk_mode == 2) {
    // ---- MT19937 2-way walks of the comparator (aesc_baseline.h:363-379):
    // position p's walks consume draws mt_off[p] .. in the order the single
    // global generator hands them out (walk i, hop len: the walker from src,
    // then the walker from v); each position's differences are added in one
    // float chain, so a lane takes a whole position
    for (uint32_t p = r0 + ln; p < r1; p += 64) {
      const int rest = A.taus[p] - ell;
      if (rest <= 0) continue;
      const uint64_t nr = A.nr_pos[p];
      const uint32_t v = A.adj[p];
      const uint32_t* d = A.mt_draws + A.mt_off[p];
      float x = 0.0f;
      for (uint64_t i = 0; i < nr; ++i) {
        uint32_t c1 = src, c2 = v;
        for (int len = 0; len < rest; ++len) {
          const uint2 d1 = A.rowd[c1];
          c1 = A.adj[d1.x + d[0] % d1.y];
          const uint2 d2 = A.rowd[c2];
          c2 = A.adj[d2.x + d[1] % d2.y];
          d += 2;
          x += (pvec[c1] - pvec[c2]);
        }
      }
      A.preds[p] += x / (ds * static_cast<float>(nr));
      if (A.tr_x) A.tr_x[p] = x;
    }
  }

  // ---- 2-way SABA walks (aesc.h:241-391), eight positions per wave
  const uint32_t grp = ln >> 3, b = ln & 7u;
  for (uint32_t pb = r0; pb < r1 && A.walk_mode == 0; pb += 8) {
    const uint32_t p = pb + grp;
    uint64_t nr = 0;
    uint32_t L = 0, v = 0;
    if (p < r1) {
      const int rest = A.taus[p] - ell;
      if (rest > 0) {
        nr = A.nr_pos[p];
        L = static_cast<uint32_t>(rest);
        v = A.adj[p];
      }
    }
    const uint64_t nr_max = wave_max_u64(nr);
    if (nr_max == 0) continue;
    uint32_t L_max = L;
    for (int off = 32; off > 0; off >>= 1) L_max = max(L_max, (uint32_t)__shfl_xor(L_max, off));
    const uint32_t step = nr ? static_cast<uint32_t>(0xFFFFFFFFull / nr) : 0u;
    float xb = 0.0f;
    for (uint64_t i = 0; i < nr_max; i += 8) {
      const bool act = i < nr;
      uint32_t c1 = src, c2 = v;
      uint32_t s1 = (static_cast<uint32_t>(i) + b + 1u) * step;
      uint32_t s2 = s1, st = s1;
      bool fb = false;
      for (uint32_t h = 0; h < L_max; ++h) {
        const bool live = act && h < L;
        const uint64_t z = __ballot(live && !fb && (s1 == 0u || s2 == 0u));
        fb = fb || (((z >> (grp * 8u)) & 0xFFull) != 0ull);
        if (live) {
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
          xb = xb + (pvec[c1] - pvec[c2]);
        }
      }
    }
    float sum = 0.0f;
    for (uint32_t k = 0; k < 8; ++k) sum = sum + __shfl(xb, static_cast<int>(grp * 8u + k));
    if (b == 0 && nr) {
      A.preds[p] += sum / (ds * static_cast<float>(nr));
      if (A.tr_x) A.tr_x[p] = sum;
    }
  }
  if (ln == 0) {
    if (A.tr_ell) A.tr_ell[src] = ell;
    if (A.tr_emax) A.tr_emax[src] = emax;
    if (A.tr_gmax) A.tr_gmax[src] = gmax;
    if (A.tr_gmin) A.tr_gmin[src] = gmin;
    if (A.tr_qsize) A.tr_qsize[src] = qsize;
  }
  __syncthreads();
}

template <bool kLds>
__global__ void __launch_bounds__(64) source_faithful_kernel(SrcArgs A) {
  extern __shared__ float lds_pvec[];
  const uint32_t n = A.n;
  float* pvec;
  uint8_t* seen;
  if constexpr (kLds) {
    pvec = lds_pvec;
    seen = reinterpret_cast<uint8_t*>(lds_pvec + n);
  } else {
    pvec = A.pvec_g + static_cast<uint64_t>(blockIdx.x) * n;
    seen = A.seen_g + static_cast<uint64_t>(blockIdx.x) * n;
  }
  // two queues of n entries, each followed by 64 spare words
  uint32_t* qa = A.qbuf + static_cast<uint64_t>(blockIdx.x) * (2ull * n + 128u);
  uint32_t* qb = qa + n + 64u;
  for (uint32_t i = lane_id(); i < n; i += 64) seen[i] = 0;
  __syncthreads();
  const uint32_t count = A.src_list ? A.n_src : n;
  for (;;) {
    uint32_t k = 0;
    if (lane_id() == 0) k = atomicAdd(A.next_src, 1u);
    k = __shfl(k, 0);
    if (k >= count) break;
    faithful_source<kLds>(A, A.src_list ? A.src_list[k] : k, pvec, seen, qa, qb);
  }
}

__global__ void __launch_bounds__(256)
symmetrize_kernel(const float* __restrict__ preds, const int32_t* __restrict__ src_of, uint64_t m,
                  float* __restrict__ out) {
  for (uint64_t t = blockIdx.x * 256ull + threadIdx.x; t < m; t += gridDim.x * 256ull) {
    const int32_t s = src_of[t];
    out[t] = preds[t] + (s >= 0 ? preds[s] : 0.0f);
  }
}

__global__ void fill_i32_kernel(int32_t* p, uint64_t n, int32_t v) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull) p[i] = v;
}
__global__ void fill_words_kernel(uint32_t* p, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

uint32_t grid_for(uint64_t work, uint32_t block, uint32_t cap) {
  const uint64_t g = (work + block - 1) / block;
  return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(g, cap)));
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Largest initial bound tau0 (aesc.h:188) per graph of a union CSR, plus 2
// where ceil() is within rounding distance: sizes the refinement tables.
__global__ void __launch_bounds__(256)
tau0_per_graph_kernel(const uint32_t* __restrict__ idx, const uint32_t* __restrict__ adj,
                      uint64_t m, uint32_t n, const uint32_t* __restrict__ voff, uint32_t n_graphs,
                      const double* __restrict__ gp, uint32_t* g_tau0max) {
  for (uint64_t p = blockIdx.x * 256ull + threadIdx.x; p < m; p += gridDim.x * 256ull) {
    const uint32_t u = row_of(idx, n, p);
    const uint32_t v = adj[p];
    if (v < u) continue;
    uint32_t lo = 0, hi = n_graphs;
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (voff[mid] <= u)
        lo = mid;
      else
        hi = mid;
    }
    bool near = false;
    const double du = idx[v + 1] - idx[v], dv = idx[u + 1] - idx[u];
    const int t0 = tau_bound_dev(du, dv, static_cast<float>(gp[3 * lo]), gp[3 * lo + 1],
                                 gp[3 * lo + 2], 0.0f, 0.0f, near);
    atomicMax(&g_tau0max[lo], static_cast<uint32_t>(t0) + (near ? 2u : 0u));
  }
}

// Device-side state attached to a dgraph for tgtp (host copies for fixups).
}  // namespace

double tgtp_now_ms() { return now_ms(); }

int launch_fill_i32(tgtp_ctx* ctx, int32_t* p, uint64_t n, int32_t v) {
  fill_i32_kernel<<<grid_for(n, 256, 4096), 256, 0, ctx->stream>>>(p, n, v);
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  return TGTP_OK;
}

int launch_fill_words(tgtp_ctx* ctx, uint32_t* p, uint32_t n) {
  fill_words_kernel<<<1, 64, 0, ctx->stream>>>(p, n);
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  return TGTP_OK;
}

int launch_tau0_per_graph(tgtp_ctx* ctx, const uint32_t* idx, const uint32_t* adj, uint64_t m,
                          uint32_t n, const uint32_t* voff, uint32_t n_graphs, const double* gp,
                          uint32_t* g_tau0max) {
  tau0_per_graph_kernel<<<grid_for(m, 256, static_cast<uint32_t>(ctx->num_cus) * 16), 256, 0,
                          ctx->stream>>>(idx, adj, m, n, voff, n_graphs, gp, g_tau0max);
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  return TGTP_OK;
}

// Internal: fill host copies + symmetrization map once per dgraph.
int tgtp_ensure_prep(tgtp_ctx* ctx, tgtp_dgraph* dg) {
  if (dg->prep) return TGTP_OK;
  auto* P = new (std::nothrow) TgtpPrep();
  if (!P) return set_error(ctx, TGTP_E_NOMEM, "host allocation");
  P->h_idx.resize(dg->n + 1);
  P->h_adj.resize(dg->m);
  hipStream_t s = ctx->stream;
  TGTP_HIP_CHECK(ctx, hipMemcpyAsync(P->h_idx.data(), dg->idx, (dg->n + 1) * 4,
                                     hipMemcpyDeviceToHost, s));
  if (dg->m)
    TGTP_HIP_CHECK(ctx, hipMemcpyAsync(P->h_adj.data(), dg->adj, dg->m * 4,
                                       hipMemcpyDeviceToHost, s));
  TGTP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  std::vector<int32_t> src;
  const int shape = tgtp_host::symmetrize_sources(P->h_idx.data(), P->h_adj.data(), dg->n, dg->m, src);
  if (shape) {
    delete P;
    return set_error(ctx, TGTP_E_ARG,
                     shape == 1 ? "graph rows must be sorted"
                                : "graph must store both directions of every edge (no reverse of some (u, v))");
  }
  if (hipMalloc(&P->d_src_of, std::max<uint64_t>(dg->m, 1) * 4) != hipSuccess) {
    delete P;
    return set_error(ctx, TGTP_E_NOMEM, "device allocation: symmetrization map");
  }
  TGTP_HIP_CHECK(ctx, hipMemcpyAsync(P->d_src_of, src.data(), dg->m * 4, hipMemcpyHostToDevice, s));
  TGTP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  dg->prep = P;
  dg->prep_free = [](void* p) {
    auto* q = static_cast<TgtpPrep*>(p);
    if (q->d_src_of) (void)hipFree(q->d_src_of);
    if (q->fast_plan && q->fast_plan_free) q->fast_plan_free(q->fast_plan);
    delete q;
  };
  return TGTP_OK;
}

int tgtp_ensure_eigens(tgtp_ctx* ctx, tgtp_dgraph* dg, const tgtp_eigens* e) {
  if (!e || !e->values || !e->vectors || e->omega < 2 || static_cast<uint64_t>(e->omega) > dg->n)
    return set_error(ctx, TGTP_E_ARG, "omega must be between 2 and vertex count");
  if (dg->eig_key == e->vectors && dg->eig_vals == e->values && dg->eig_omega == e->omega &&
      dg->X)
    return TGTP_OK;
  // forget the old key first: a failed upload must not leave it matching
  dg->eig_key = nullptr;
  dg->eig_vals = nullptr;
  if (dg->X) (void)hipFree(dg->X);
  if (dg->lam) (void)hipFree(dg->lam);
  dg->X = nullptr;
  dg->lam = nullptr;
  const uint64_t nx = dg->n * static_cast<uint64_t>(e->omega);
  if (hipMalloc(&dg->X, nx * 4) != hipSuccess || hipMalloc(&dg->lam, e->omega * 8) != hipSuccess)
    return set_error(ctx, TGTP_E_NOMEM, "device allocation: eigenpairs");
  TGTP_HIP_CHECK(ctx, hipMemcpyAsync(dg->X, e->vectors, nx * 4, hipMemcpyHostToDevice, ctx->stream));
  TGTP_HIP_CHECK(ctx, hipMemcpyAsync(dg->lam, e->values, e->omega * 8, hipMemcpyHostToDevice,
                                     ctx->stream));
  TGTP_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  dg->eig_key = e->vectors;
  dg->eig_vals = e->values;
  dg->eig_omega = e->omega;
  ++dg->eig_gen;
  return TGTP_OK;
}

// tau for every CSR position into d_taus (device) -- faithful arithmetic.
int tgtp_run_taus_faithful(tgtp_ctx* ctx, tgtp_dgraph* dg, const tgtp_eigens* e, float eps_half,
                             int32_t* d_taus, uint64_t* n_fixups) {
  hipStream_t s = ctx->stream;
  auto* P = static_cast<TgtpPrep*>(dg->prep);
  const int om = e->omega;
  std::vector<double> onep(om), onem(om);
  for (int i = 0; i < om; ++i) {
    onep[i] = 1 + e->values[i];
    onem[i] = 1 - e->values[i];
  }
  auto* d_misc = static_cast<char*>(ctx_buffer(ctx, kBufMisc, 2 * om * 8 + 4096));
  auto* d_words = static_cast<uint32_t*>(ctx_buffer(ctx, kBufFlags, 64));
  auto* d_flags = static_cast<uint32_t*>(ctx_buffer(ctx, kBufStage, std::max<uint64_t>(dg->m, 1) * 4));
  if (!d_misc || !d_words || !d_flags) return set_error(ctx, TGTP_E_NOMEM, "tau scratch");
  double* d_onep = reinterpret_cast<double*>(d_misc);
  double* d_onem = d_onep + om;
  TGTP_HIP_CHECK(ctx, hipMemcpyAsync(d_onep, onep.data(), om * 8, hipMemcpyHostToDevice, s));
  TGTP_HIP_CHECK(ctx, hipMemcpyAsync(d_onem, onem.data(), om * 8, hipMemcpyHostToDevice, s));
  fill_words_kernel<<<1, 64, 0, s>>>(d_words, 16);
  TauArgs A{};
  A.idx = dg->idx;
  A.adj = dg->adj;
  A.X = dg->X;
  A.onep = d_onep;
  A.onem = d_onem;
  A.m = dg->m;
  A.n = static_cast<uint32_t>(dg->n);
  A.omega = om;
  A.eps = eps_half;
  const double l1 = e->values[1], lw = e->values[om - 1];
  A.sq1m_1 = 1.0 - (l1 * l1);
  A.b_1 = std::log(1.0 / std::fabs(l1));
  A.sq1m_w = 1.0 - (lw * lw);
  A.b_w = std::log(1.0 / std::fabs(lw));
  A.taus = d_taus;
  A.flags = d_flags;
  A.n_flags = d_words;
  A.tau0_max = d_words + 1;
  const uint32_t cap = static_cast<uint32_t>(ctx->num_cus) * 16;
  tau0_kernel<<<grid_for(dg->m, 256, cap), 256, 0, s>>>(A);
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  uint32_t words[2] = {0, 0};
  TGTP_HIP_CHECK(ctx, hipMemcpyAsync(words, d_words, 8, hipMemcpyDeviceToHost, s));
  TGTP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  const uint32_t k_rows = std::min<uint32_t>(words[1] / 2 + 3, 1u << 20);
  std::vector<double> pw = tgtp_host::power_table(e->values, om, k_rows);
  auto* d_pw = static_cast<double*>(ctx_buffer(ctx, kBufEdgeInfo, pw.size() * 8));
  if (!d_pw) return set_error(ctx, TGTP_E_NOMEM, "tau power table");
  TGTP_HIP_CHECK(ctx, hipMemcpyAsync(d_pw, pw.data(), pw.size() * 8, hipMemcpyHostToDevice, s));
  A.pw = d_pw;
  A.k_rows = k_rows;
  fill_i32_kernel<<<grid_for(dg->m, 256, cap), 256, 0, s>>>(d_taus, dg->m, 0);
  taus_faithful_kernel<<<grid_for(dg->m, 256, cap), 256, 0, s>>>(A);
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  TGTP_HIP_CHECK(ctx, hipMemcpyAsync(words, d_words, 4, hipMemcpyDeviceToHost, s));
  TGTP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  *n_fixups = words[0];
  if (words[0]) {
    // recompute the flagged edges with the C library's log
    std::vector<uint32_t> flags(words[0]);
    TGTP_HIP_CHECK(ctx, hipMemcpyAsync(flags.data(), d_flags, flags.size() * 4,
                                       hipMemcpyDeviceToHost, s));
    std::vector<int32_t> taus(dg->m);
    TGTP_HIP_CHECK(ctx, hipMemcpyAsync(taus.data(), d_taus, dg->m * 4, hipMemcpyDeviceToHost, s));
    TGTP_HIP_CHECK(ctx, hipStreamSynchronize(s));
    const uint32_t* idx = P->h_idx.data();
    const uint32_t* adj = P->h_adj.data();
    for (uint32_t p : flags) {
      const uint32_t u = static_cast<uint32_t>(
          std::upper_bound(idx, idx + dg->n + 1, p) - idx - 1);
      const uint32_t v = adj[p];
      const int tau = tgtp_host::edge_tau(idx, e->vectors, om, e->values, dg->m, eps_half, u, v);
      taus[p] = tau;
      taus[std::lower_bound(adj + idx[v], adj + idx[v + 1], u) - adj] = tau;
    }
    TGTP_HIP_CHECK(ctx, hipMemcpyAsync(d_taus, taus.data(), dg->m * 4, hipMemcpyHostToDevice, s));
    TGTP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  }
  return TGTP_OK;
}

// sl (optional): run only the listed sources and also return every row.
struct SourceList {
  const uint32_t* d_list;
  uint32_t count;
  float* rows_host;  // [m] preds before symmetrization
};

static int run_score_faithful(tgtp_ctx* ctx, tgtp_dgraph* dg, const tgtp_eigens* e,
                              const tgtp_params* prm, float* scores_out, tgtp_stats* stats,
                              const TraceOut* tr, const SourceList* sl = nullptr) {
  const double t_host0 = now_ms();
  hipStream_t s = ctx->stream;
  const uint64_t n = dg->n, m = dg->m;
  auto* P = static_cast<TgtpPrep*>(dg->prep);
  auto* d_taus = static_cast<int32_t*>(ctx_buffer(ctx, kBufTaus, std::max<uint64_t>(m, 1) * 4));
  auto* d_preds = static_cast<float*>(ctx_buffer(ctx, kBufPreds, std::max<uint64_t>(m, 1) * 8));
  auto* d_nr = static_cast<uint64_t*>(ctx_buffer(ctx, kBufSrcInfo, std::max<uint64_t>(m, 1) * 8));
  if (!d_taus || !d_preds || !d_nr) return set_error(ctx, TGTP_E_NOMEM, "tgtp buffers");
  float* d_out = d_preds + m;
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[4], s));
  uint64_t fixups = 0;
  int rc = tgtp_run_taus_faithful(ctx, dg, e, static_cast<float>(prm->epsilon / 2.0), d_taus, &fixups);
  if (rc) return rc;
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[5], s));

  // per-source kernel configuration
  const size_t lds_bytes = n * 5 + 16;
  const bool use_lds = lds_bytes <= 64 * 1024;
  uint32_t per_cu = use_lds ? static_cast<uint32_t>(std::min<size_t>(8, (160 * 1024) / lds_bytes))
                            : 4u;
  per_cu = std::max<uint32_t>(per_cu, 1);
  const uint64_t n_run = sl ? std::max<uint32_t>(sl->count, 1u) : n;
  const uint32_t slots =
      static_cast<uint32_t>(std::min<uint64_t>(n_run, static_cast<uint64_t>(ctx->num_cus) * per_cu));
  auto* d_q = static_cast<uint32_t*>(
      ctx_buffer(ctx, kBufQueue, static_cast<uint64_t>(slots) * (2 * n + 128) * 4));
  auto* d_words = static_cast<uint32_t*>(ctx_buffer(ctx, kBufSums, 64));
  float* d_pv = nullptr;
  if (!use_lds) {
    d_pv = static_cast<float*>(ctx_buffer(ctx, kBufPvec, static_cast<uint64_t>(slots) * n * 5));
    if (!d_pv) return set_error(ctx, TGTP_E_NOMEM, "per-source state");
  }
  if (!d_q || !d_words) return set_error(ctx, TGTP_E_NOMEM, "frontier queues");
  fill_words_kernel<<<1, 64, 0, s>>>(d_words, 16);
  SrcArgs A{};
  A.rowd = dg->rowd;
  A.adj = dg->adj;
  A.degf = dg->degf;
  A.taus = d_taus;
  A.preds = d_preds;
  A.nr_pos = d_nr;
  A.qbuf = d_q;
  A.pvec_g = d_pv;
  A.seen_g = d_pv ? reinterpret_cast<uint8_t*>(d_pv + static_cast<uint64_t>(slots) * n) : nullptr;
  A.next_src = d_words;
  A.n_ambiguous = d_words + 1;
  A.status = d_words + 2;
  A.src_list = sl ? sl->d_list : nullptr;
  A.n_src = sl ? sl->count : 0u;
  A.n = static_cast<uint32_t>(n);
  A.gamma = prm->gamma;
  A.eps = prm->epsilon;
  A.logmd = std::log(m / prm->delta);
  float* tr_buf = nullptr;
  if (tr) {
    tr_buf = static_cast<float*>(ctx_buffer(ctx, kBufTrace, (5 * n + m) * 4 + 64));
    if (!tr_buf) return set_error(ctx, TGTP_E_NOMEM, "trace buffers");
    A.tr_ell = reinterpret_cast<int32_t*>(tr_buf);
    A.tr_emax = tr_buf + n;
    A.tr_gmax = tr_buf + 2 * n;
    A.tr_gmin = tr_buf + 3 * n;
    A.tr_qsize = reinterpret_cast<uint32_t*>(tr_buf + 4 * n);
    A.tr_x = tr_buf + 5 * n;
    fill_i32_kernel<<<grid_for(5 * n + m, 256, 4096), 256, 0, s>>>(
        reinterpret_cast<int32_t*>(tr_buf), 5 * n + m, 0);
  }
  auto launch_sources = [&]() -> int {
    if (use_lds)
      source_faithful_kernel<true><<<slots, 64, lds_bytes, s>>>(A);
    else
      source_faithful_kernel<false><<<slots, 64, 0, s>>>(A);
    TGTP_HIP_CHECK(ctx, hipGetLastError());
    return TGTP_OK;
  };
  if (prm->mode & TGTP_WALK2_MT19937) {
    // the comparator's walks draw from one generator in source and position
    // order (aesc_baseline.h:261, 338): a counting pass gives every
    // position's walk count and rest, their prefix sums give each
    // position's offset in the stream, which is generated once; the walk
    // pass then repeats the (deterministic) per-source pipeline with walks
    auto* d_ell = static_cast<int32_t*>(ctx_buffer(ctx, kBufMtEll, n * 4 + 64));
    if (!d_ell) return set_error(ctx, TGTP_E_NOMEM, "walk rounds buffer");
    int32_t* keep_ell = A.tr_ell;
    A.tr_ell = keep_ell ? keep_ell : d_ell;
    A.walk_mode = 1;
    if ((rc = launch_sources())) return rc;
    std::vector<uint64_t> nr_h(m), off(m + 1, 0);
    std::vector<int32_t> tau_h(m), ell_h(n);
    TGTP_HIP_CHECK(ctx, hipMemcpyAsync(nr_h.data(), d_nr, m * 8, hipMemcpyDeviceToHost, s));
    TGTP_HIP_CHECK(ctx, hipMemcpyAsync(tau_h.data(), d_taus, m * 4, hipMemcpyDeviceToHost, s));
    TGTP_HIP_CHECK(ctx, hipMemcpyAsync(ell_h.data(), A.tr_ell, n * 4, hipMemcpyDeviceToHost, s));
    uint32_t w0[3] = {0, 0, 0};
    TGTP_HIP_CHECK(ctx, hipMemcpyAsync(w0, d_words, 12, hipMemcpyDeviceToHost, s));
    TGTP_HIP_CHECK(ctx, hipStreamSynchronize(s));
    if (w0[2] & 1u)
      return set_error(ctx, TGTP_E_UNSUPPORTED,
                       "a walk count exceeded 2^36 walks per edge; scores are not computed");
    const uint32_t* hidx = P->h_idx.data();
    for (uint64_t u = 0; u < n; ++u)
      for (uint32_t p = hidx[u]; p < hidx[u + 1]; ++p) {
        const int rest = tau_h[p] - ell_h[u];
        off[p + 1] = off[p] + (rest > 0 ? 2ull * nr_h[p] * static_cast<uint64_t>(rest) : 0ull);
      }
    const uint64_t draws = off[m];
    if (draws > (1ull << 32))
      return set_error(ctx, TGTP_E_UNSUPPORTED,
                       "the MT19937 walks need more than 2^32 draws (16 GiB of stream)");
    auto* d_off = static_cast<uint64_t*>(ctx_buffer(ctx, kBufMtOff, (m + 1) * 8));
    auto* d_draws = static_cast<uint32_t*>(ctx_buffer(ctx, kBufMtDraws, draws * 4 + 64));
    if (!d_off || !d_draws) return set_error(ctx, TGTP_E_NOMEM, "MT19937 stream");
    TGTP_HIP_CHECK(ctx, hipMemcpyAsync(d_off, off.data(), (m + 1) * 8, hipMemcpyHostToDevice, s));
    if ((rc = launch_mt_stream(ctx, 42u, draws, d_draws))) return rc;  // rndgen.seed(42)
    fill_words_kernel<<<1, 64, 0, s>>>(d_words, 16);
    A.walk_mode = 2;
    A.mt_draws = d_draws;
    A.mt_off = d_off;
    A.tr_ell = keep_ell;
    TGTP_HIP_CHECK(ctx, hipStreamSynchronize(s));  // host vectors above stay alive until here
  }
  if ((rc = launch_sources())) return rc;
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[6], s));
  symmetrize_kernel<<<grid_for(m, 256, 4096), 256, 0, s>>>(d_preds, P->d_src_of, m, d_out);
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  TGTP_HIP_CHECK(ctx, hipEventRecord(ctx->ev[7], s));
  TGTP_HIP_CHECK(ctx, hipMemcpyAsync(scores_out, d_out, m * 4, hipMemcpyDeviceToHost, s));
  if (sl) TGTP_HIP_CHECK(ctx, hipMemcpyAsync(sl->rows_host, d_preds, m * 4, hipMemcpyDeviceToHost, s));
  uint32_t words[3] = {0, 0, 0};
  TGTP_HIP_CHECK(ctx, hipMemcpyAsync(words, d_words, 12, hipMemcpyDeviceToHost, s));
  if (tr) {
    if (tr->taus) TGTP_HIP_CHECK(ctx, hipMemcpyAsync(tr->taus, d_taus, m * 4, hipMemcpyDeviceToHost, s));
    if (tr->nr) TGTP_HIP_CHECK(ctx, hipMemcpyAsync(tr->nr, d_nr, m * 8, hipMemcpyDeviceToHost, s));
    if (tr->preds_row)
      TGTP_HIP_CHECK(ctx, hipMemcpyAsync(tr->preds_row, d_preds, m * 4, hipMemcpyDeviceToHost, s));
    if (tr->walk_sum)
      TGTP_HIP_CHECK(ctx, hipMemcpyAsync(tr->walk_sum, A.tr_x, m * 4, hipMemcpyDeviceToHost, s));
    if (tr->ell) TGTP_HIP_CHECK(ctx, hipMemcpyAsync(tr->ell, A.tr_ell, n * 4, hipMemcpyDeviceToHost, s));
    if (tr->edge_max)
      TGTP_HIP_CHECK(ctx, hipMemcpyAsync(tr->edge_max, A.tr_emax, n * 4, hipMemcpyDeviceToHost, s));
    if (tr->gmax) TGTP_HIP_CHECK(ctx, hipMemcpyAsync(tr->gmax, A.tr_gmax, n * 4, hipMemcpyDeviceToHost, s));
    if (tr->gmin) TGTP_HIP_CHECK(ctx, hipMemcpyAsync(tr->gmin, A.tr_gmin, n * 4, hipMemcpyDeviceToHost, s));
    if (tr->q_size)
      TGTP_HIP_CHECK(ctx, hipMemcpyAsync(tr->q_size, A.tr_qsize, n * 4, hipMemcpyDeviceToHost, s));
  }
  TGTP_HIP_CHECK(ctx, hipStreamSynchronize(s));
  if (words[2] & 1u)
    return set_error(ctx, TGTP_E_UNSUPPORTED,
                     "a walk count exceeded 2^36 walks per edge; scores are not computed");
  if (words[2] & 2u)
    return set_error(ctx, TGTP_E_ARG, "non-finite diffusion values (check the eigenpairs)");
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    float a = 0, b = 0, c = 0;
    TGTP_HIP_CHECK(ctx, hipEventElapsedTime(&a, ctx->ev[4], ctx->ev[7]));
    TGTP_HIP_CHECK(ctx, hipEventElapsedTime(&b, ctx->ev[4], ctx->ev[5]));
    TGTP_HIP_CHECK(ctx, hipEventElapsedTime(&c, ctx->ev[5], ctx->ev[6]));
    stats->kernel_ms = a;
    stats->taus_ms = b;
    stats->diffuse_ms = c;
    stats->host_fixups = fixups;
    stats->work_items = words[1];  // sources whose top-gamma needed the exact permutation
    stats->e2e_ms = now_ms() - t_host0;
  }
  return TGTP_OK;
}

static int validate_params(tgtp_ctx* ctx, const tgtp_params* p) {
  if (!p) return set_error(ctx, TGTP_E_ARG, "null params");
  if (!(p->epsilon > 0 && p->epsilon < 1) || p->gamma < 1 || !(p->delta > 0))
    return set_error(ctx, TGTP_E_ARG,
                     "require 0 < epsilon < 1, integer omega >= 2, integer gamma >= 1, delta > 0");
  if (!tgtp_mode_valid(p->mode)) return set_error(ctx, TGTP_E_ARG, "unknown tgtp mode");
  return TGTP_OK;
}

static int run_score(tgtp_ctx* ctx, tgtp_dgraph* dg, const tgtp_eigens* e, const tgtp_params* p,
                     float* out, tgtp_stats* stats, const TraceOut* tr) {
  if (tgtp_mode_base(p->mode) == TGTP_MODE_FAITHFUL)
    return run_score_faithful(ctx, dg, e, p, out, stats, tr);
  return tgtp_run_score_fast(ctx, dg, e, p, out, stats, tr);
}

int tgtp_check_degrees(tgtp_ctx* ctx, tgtp_dgraph* dg) {
  auto* P = static_cast<TgtpPrep*>(dg->prep);
  for (uint64_t v = 0; v < dg->n; ++v)
    if (P->h_idx[v + 1] == P->h_idx[v])
      return set_error(ctx, TGTP_E_ARG,
                       "graph vertex IDs must be contiguous with no isolated vertices");
  return TGTP_OK;
}

extern "C" int tgtp_score_dev(tgtp_ctx* ctx, const tgtp_dgraph* cdg, const tgtp_eigens* e,
                              const tgtp_params* p, float* scores_out, tgtp_stats* stats) {
  if (!ctx) return TGTP_E_ARG;
  if (!cdg || !scores_out) return set_error(ctx, TGTP_E_ARG, "null graph or output");
  auto* dg = const_cast<tgtp_dgraph*>(cdg);
  int rc = validate_params(ctx, p);
  if (rc) return rc;
  TGTP_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if ((rc = tgtp_ensure_prep(ctx, dg))) return rc;
  if ((rc = tgtp_check_degrees(ctx, dg))) return rc;
  if ((rc = tgtp_ensure_eigens(ctx, dg, e))) return rc;
  return run_score(ctx, dg, e, p, scores_out, stats, nullptr);
}

extern "C" int tgtp_score_sources(tgtp_ctx* ctx, const tgtp_dgraph* cdg, const tgtp_eigens* e,
                                  const tgtp_params* p, uint32_t n_ranges,
                                  const uint32_t* ranges, float* rows_out, float* scores_out,
                                  tgtp_stats* stats) {
  if (!ctx) return TGTP_E_ARG;
  if (!cdg || !ranges || n_ranges == 0)
    return set_error(ctx, TGTP_E_ARG, "null graph or empty source ranges");
  auto* dg = const_cast<tgtp_dgraph*>(cdg);
  int rc = validate_params(ctx, p);
  if (rc) return rc;
  if (p->mode & TGTP_WALK2_MT19937)  // its stream offsets depend on every earlier source
    return set_error(ctx, TGTP_E_ARG, "source subsets do not take the MT19937 walks");
  std::vector<std::pair<uint32_t, uint32_t>> rg(n_ranges);
  uint64_t count = 0;
  for (uint32_t k = 0; k < n_ranges; ++k) {
    rg[k] = {ranges[2 * k], ranges[2 * k + 1]};
    if (rg[k].first >= rg[k].second || rg[k].second > dg->n ||
        (k && rg[k].first < rg[k - 1].second))
      return set_error(ctx, TGTP_E_ARG,
                       "source ranges must be non-empty, ascending, disjoint and within [0, n)");
    count += rg[k].second - rg[k].first;
  }
  TGTP_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if ((rc = tgtp_ensure_prep(ctx, dg))) return rc;
  if ((rc = tgtp_check_degrees(ctx, dg))) return rc;
  if ((rc = tgtp_ensure_eigens(ctx, dg, e))) return rc;
  const uint64_t m = dg->m;
  std::vector<float> all_scores(m), all_rows(m);
  if (tgtp_mode_base(p->mode) == TGTP_MODE_FAITHFUL) {
    std::vector<uint32_t> list;
    list.reserve(count);
    for (const auto& r : rg)
      for (uint32_t v = r.first; v < r.second; ++v) list.push_back(v);
    auto* d_list = static_cast<uint32_t*>(ctx_buffer(ctx, kBufSubset, count * 4));
    if (!d_list) return set_error(ctx, TGTP_E_NOMEM, "source list");
    TGTP_HIP_CHECK(ctx, hipMemcpyAsync(d_list, list.data(), count * 4, hipMemcpyHostToDevice,
                                       ctx->stream));
    const SourceList sl{d_list, static_cast<uint32_t>(count), all_rows.data()};
    rc = run_score_faithful(ctx, dg, e, p, all_scores.data(), stats, nullptr, &sl);
  } else {
    rc = tgtp_run_score_fast_sources(ctx, dg, e, p, rg, all_scores.data(), all_rows.data(), stats);
  }
  if (rc) return rc;
  // keep rows of listed sources, scores of positions whose ends are both listed
  const auto* P = static_cast<TgtpPrep*>(dg->prep);
  auto listed = [&](uint32_t v) {
    auto it = std::upper_bound(rg.begin(), rg.end(), v,
                               [](uint32_t x, const std::pair<uint32_t, uint32_t>& r) {
                                 return x < r.first;
                               });
    return it != rg.begin() && v < std::prev(it)->second;
  };
  for (const auto& r : rg)
    for (uint32_t u = r.first; u < r.second; ++u)
      for (uint32_t pos = P->h_idx[u]; pos < P->h_idx[u + 1]; ++pos) {
        if (rows_out) rows_out[pos] = all_rows[pos];
        if (scores_out && listed(P->h_adj[pos])) scores_out[pos] = all_scores[pos];
      }
  if (stats) stats->work_items = count;
  return TGTP_OK;
}

extern "C" int tgtp_score(tgtp_ctx* ctx, const tgtp_graph* g, const tgtp_eigens* e,
                          const tgtp_params* p, float* scores_out, tgtp_stats* stats) {
  if (!ctx) return TGTP_E_ARG;
  const double t0 = now_ms();
  tgtp_dgraph* dg = nullptr;
  int rc = tgtp_graph_upload(ctx, g, &dg);
  if (rc) return rc;
  rc = tgtp_score_dev(ctx, dg, e, p, scores_out, stats);
  tgtp_graph_free(ctx, dg);
  if (!rc && stats) stats->e2e_ms = now_ms() - t0;
  return rc;
}

extern "C" int tgtp_score_trace(tgtp_ctx* ctx, const tgtp_dgraph* cdg, const tgtp_eigens* e,
                                const tgtp_params* p, float* scores_out, tgtp_stats* stats,
                                int32_t* taus, uint64_t* nr, float* walk_sum, float* preds_row,
                                int32_t* ell, float* edge_max, float* gmax, float* gmin,
                                uint32_t* q_size) {
  if (!ctx) return TGTP_E_ARG;
  if (!cdg || !scores_out) return set_error(ctx, TGTP_E_ARG, "null graph or output");
  auto* dg = const_cast<tgtp_dgraph*>(cdg);
  int rc = validate_params(ctx, p);
  if (rc) return rc;
  TGTP_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if ((rc = tgtp_ensure_prep(ctx, dg))) return rc;
  if ((rc = tgtp_check_degrees(ctx, dg))) return rc;
  if ((rc = tgtp_ensure_eigens(ctx, dg, e))) return rc;
  TraceOut tr{taus, nr, walk_sum, preds_row, ell, edge_max, gmax, gmin, q_size};
  return run_score(ctx, dg, e, p, scores_out, stats, &tr);
}

extern "C" int tgtp_taus(tgtp_ctx* ctx, const tgtp_graph* g, const tgtp_eigens* e,
                         float eps_half, int32_t mode, int32_t* taus_out, tgtp_stats* stats) {
  if (!ctx) return TGTP_E_ARG;
  if (!taus_out) return set_error(ctx, TGTP_E_ARG, "null output");
  if (!tgtp_mode_valid(mode)) return set_error(ctx, TGTP_E_ARG, "unknown tgtp mode");
  if (!(eps_half > 0)) return set_error(ctx, TGTP_E_ARG, "epsilon must be positive");
  const double t0 = now_ms();
  tgtp_dgraph* dg = nullptr;
  int rc = tgtp_graph_upload(ctx, g, &dg);
  if (rc) return rc;
  auto cleanup = [&](int r) {
    tgtp_graph_free(ctx, dg);
    return r;
  };
  if ((rc = tgtp_ensure_prep(ctx, dg))) return cleanup(rc);
  if ((rc = tgtp_ensure_eigens(ctx, dg, e))) return cleanup(rc);
  auto* d_taus = static_cast<int32_t*>(ctx_buffer(ctx, kBufTaus, std::max<uint64_t>(dg->m, 1) * 4));
  if (!d_taus) return cleanup(set_error(ctx, TGTP_E_NOMEM, "taus buffer"));
  hipStream_t s = ctx->stream;
  if (hipEventRecord(ctx->ev[4], s) != hipSuccess) return cleanup(TGTP_E_HIP);
  uint64_t fixups = 0;
  rc = tgtp_mode_base(mode) == TGTP_MODE_FAITHFUL ? tgtp_run_taus_faithful(ctx, dg, e, eps_half, d_taus, &fixups)
                                 : tgtp_run_taus_fast(ctx, dg, e, eps_half, mode, d_taus);
  if (rc) return cleanup(rc);
  if (hipEventRecord(ctx->ev[5], s) != hipSuccess ||
      hipMemcpyAsync(taus_out, d_taus, dg->m * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return cleanup(set_error(ctx, TGTP_E_HIP, "taus copy"));
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    float a = 0;
    (void)hipEventElapsedTime(&a, ctx->ev[4], ctx->ev[5]);
    stats->kernel_ms = stats->taus_ms = a;
    stats->host_fixups = fixups;
    stats->e2e_ms = now_ms() - t0;
  }
  return cleanup(TGTP_OK);
}
