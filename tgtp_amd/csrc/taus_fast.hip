// taus_fast.hip -- tau bounds (aesc.h:164-216) for FAST_FP32 / FAST_BF16 as
// an MFMA contraction.
//
// For an undirected edge (u, v) let D_i = (x_ui - x_vi)^2, i in [1, omega-2]
// (0 elsewhere).  The reference's loops are dot products of D with fixed
// columns:  upsilon = <D, 1 + lambda> / m  and, at refinement step k,
// delta_k = <D, lambda^(2(k+1)) / (1 - lambda)> / m.  Sixteen edges form the
// rows of a tile, the columns [1 + lambda | lambda^2/(1-lambda) |
// lambda^4/(1-lambda) | ...] form W, and C = D W is accumulated in FP32 by
// v_mfma_f32_16x16x4_f32 (FAST_FP32: D and W in FP32) or
// v_mfma_f32_16x16x32_bf16 (FAST_BF16: eigenvectors stored and D, W rounded
// to BF16).  The stopping test (get_tau, aesc.h:129-135) runs in FP64 on the
// tile's row after each 16-column tile; further column tiles are computed
// only while some edge of the tile is still refining.  With A.bounded a
// refinement step whose delta alone reaches eps/2 is skipped (t advances,
// tau is kept) instead of being bounded against eps/2 (aesc.h:130), so the
// kept tau bounds both the top-omega tail and the rest of the series.
//
// Lane mapping (16x16 MFMA): lane l feeds row / column l % 16 and the k
// range of group l / 16; k ranges of A and B are the same permutation of
// [0, omega_pad), so the contraction is exact in any order of k.  The result
// holds rows 4*(l/16) .. +3 of column l % 16.
#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "taus_fast.hpp"
#include "wave.hpp"

using namespace tgtp_dev;

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int kTB = 256;  // 4 waves, one 16-edge tile each per step

__device__ __forceinline__ uint32_t row_of_pos(const uint32_t* idx, uint32_t n, uint64_t p) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (idx[mid] <= p)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint32_t first_pos_of(const uint32_t* idx, const uint32_t* adj,
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

// get_tau in FP64 (aesc.h:129-135)
__device__ __forceinline__ int tau_bound64(double du, double dv, float eps, double sq1m, double b,
                                           float del, float ups) {
  const float eps_del = eps - del > 0.0 ? eps - del : eps;
  const double x = (1.0 / du + 1.0 / dv - 2.0 / du / dv - ups) / eps_del / sq1m;
  const double a = log(x < 1.0 ? 1.0 : x);
  const double c = ceil(a / b - 1);
  const int tau = static_cast<int>(c < 1.0 ? 1.0 : c);
  return tau % 2 == 0 ? tau + 1 : tau;
}

template <bool kBF16>
__global__ void __launch_bounds__(kTB) taus_fast_kernel(TauFastArgs A) {
  __shared__ float ct[kTB / 64][16][17];
  const uint32_t l = lane_id();
  const uint32_t w = wave_id();
  const uint32_t r = l & 15u, kk = l >> 4;
  const uint32_t kq = A.omega_pad / 4;  // k values per lane group
  for (uint64_t tile = static_cast<uint64_t>(blockIdx.x) * (kTB / 64) + w; tile < A.n_tiles;
       tile += static_cast<uint64_t>(gridDim.x) * (kTB / 64)) {
    const TauGraph G = A.graphs[A.tile_graph[tile]];
    const uint32_t p = A.und[tile * 16 + r];
    const bool valid = p != 0xFFFFFFFFu;
    uint32_t u = 0, v = 0;
    if (valid) {
      u = row_of_pos(A.idx, A.n, p);
      v = A.adj[p];
    }
    // D for this lane's k range, kept in registers for all column tiles
    float dreg[32];
    bf8 dbf[4];
    const uint32_t k0 = kk * kq;
    if constexpr (!kBF16) {
#pragma unroll
      for (uint32_t s = 0; s < 32; ++s) {
        const uint32_t i = k0 + s;
        float d = 0.0f;
        if (valid && s < kq && i >= 1 && i + 1 < static_cast<uint32_t>(A.omega)) {
          const float df = A.X[static_cast<uint64_t>(u) * A.omega + i] -
                           A.X[static_cast<uint64_t>(v) * A.omega + i];
          d = df * df;
        }
        dreg[s] = d;
      }
    } else {
#pragma unroll
      for (uint32_t s = 0; s < 4; ++s) {
#pragma unroll
        for (uint32_t t = 0; t < 8; ++t) {
          const uint32_t i = k0 + 8 * s + t;
          float d = 0.0f;
          if (valid && 8 * s + t < kq && i >= 1 && i + 1 < static_cast<uint32_t>(A.omega)) {
            const float df = static_cast<float>(A.Xb[static_cast<uint64_t>(u) * A.omega + i]) -
                             static_cast<float>(A.Xb[static_cast<uint64_t>(v) * A.omega + i]);
            d = df * df;
          }
          dbf[s][t] = static_cast<__bf16>(d);
        }
      }
    }
    // per-edge refinement state (lanes 0..15 own the tile's rows)
    double du = 0, dv = 0;
    float ups = 0.0f;
    int tau = 1, t = 1;
    bool done = !valid;
    if (l < 16 && valid) {
      du = static_cast<double>(A.idx[v + 1] - A.idx[v]);
      dv = static_cast<double>(A.idx[u + 1] - A.idx[u]);
    }
    for (uint32_t ctile = 0; ctile < G.n_col_tiles; ++ctile) {
      f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
      const uint64_t col = static_cast<uint64_t>(G.wt_row0) + ctile * 16 + r;
      if constexpr (!kBF16) {
        const float* wcol = A.Wt + col * A.omega_pad + k0;
#pragma unroll
        for (uint32_t s = 0; s < 32; ++s) {
          if (s < kq) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(dreg[s], wcol[s], acc, 0, 0, 0);
        }
      } else {
        const __bf16* wcol = A.Wtb + col * A.omega_pad + k0;
#pragma unroll
        for (uint32_t s = 0; s < 4; ++s) {
          if (8 * s < kq) {
            // one 16-byte load: omega_pad is a multiple of 32 and k0 of 8
            const bf8 wb = *reinterpret_cast<const bf8*>(wcol + 8 * s);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(dbf[s], wb, acc, 0, 0, 0);
          }
        }
      }
#pragma unroll
      for (uint32_t q = 0; q < 4; ++q) ct[w][4 * kk + q][r] = acc[q];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      if (l < 16 && !done) {
        uint32_t c = 0;
        if (ctile == 0) {
          ups = ct[w][l][0] / G.m;
          tau = tau_bound64(du, dv, G.eps, G.sq1m_1, G.b_1, 0.0f, 0.0f);
          c = 1;
        }
        for (; c < 16 && !done; ++c) {
          const float del = ct[w][l][c] / G.m;
          if (A.bounded && !(G.eps - del > 0.0)) {
            // the top-omega tail after t steps alone reaches eps/2, so no
            // tau >= t bounds the series with this delta; go on to the next
            // step (aesc.h:130 would bound against eps/2 alone here)
            t += 2;
            if (t >= tau) done = true;
            continue;
          }
          const int tp = tau_bound64(du, dv, G.eps, G.sq1m_w, G.b_w, del, ups);
          if (t <= tp && tp < tau) {
            tau = tp;
            t += 2;
          } else {
            done = true;
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
      if (!__any(l < 16 && !done)) break;
    }
    if (l < 16 && valid) {
      A.taus[p] = tau;
      A.taus[first_pos_of(A.idx, A.adj, v, u)] = tau;
    }
  }
}

__global__ void to_bf16_kernel(const float* __restrict__ x, uint64_t count, __bf16* __restrict__ y) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += gridDim.x * 256ull)
    y[i] = static_cast<__bf16>(x[i]);
}

}  // namespace

int launch_taus_fast(tgtp_ctx* ctx, const TauFastArgs& A, bool bf16) {
  const uint64_t blocks = std::max<uint64_t>(
      1, std::min<uint64_t>((A.n_tiles + 3) / 4, static_cast<uint64_t>(ctx->num_cus) * 8));
  if (bf16)
    taus_fast_kernel<true><<<static_cast<uint32_t>(blocks), kTB, 0, ctx->stream>>>(A);
  else
    taus_fast_kernel<false><<<static_cast<uint32_t>(blocks), kTB, 0, ctx->stream>>>(A);
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  return TGTP_OK;
}

int launch_to_bf16(tgtp_ctx* ctx, const float* x, uint64_t count, void* y) {
  const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>(8192, (count + 255) / 256));
  to_bf16_kernel<<<static_cast<uint32_t>(blocks), 256, 0, ctx->stream>>>(x, count,
                                                                       static_cast<__bf16*>(y));
  TGTP_HIP_CHECK(ctx, hipGetLastError());
  return TGTP_OK;
}
