// gather_roof.hip -- the roof of the walk engine's memory pattern: chains
// of dependent random 8-byte loads over a buffer of the walk's packed
// adjacency size (the next address depends on the loaded word, as a walk's
// next row depends on the neighbour it loaded), measured for several
// occupancies (waves per SIMD, set through an LDS reservation per workgroup)
// and several independent chains per lane.
//
//   gather_roof [--mib 160] [--hops 64] [--reps 3]
//
// Prints one line per (waves per SIMD, chains): loads/s, and GB/s at 8 B
// per load.  Every address stays inside the buffer (index mod entries).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

#define CHECK(x)                                                                      \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) {                                                           \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                    \
      std::exit(1);                                                                   \
    }                                                                                 \
  } while (0)

__device__ __forceinline__ uint32_t mix(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  return h ^ (h >> 16);
}

// Each lane runs kChains independent chains of `hops` dependent loads; the
// loaded word decides the next index.  The LDS array only sets occupancy.
template <int kChains>
__global__ void __launch_bounds__(256) chase_kernel(const uint64_t* __restrict__ buf,
                                                    uint64_t entries, uint32_t hops,
                                                    uint64_t* __restrict__ sink) {
  extern __shared__ uint32_t lds[];
  const uint32_t gid = blockIdx.x * 256u + threadIdx.x;
  uint64_t idx[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) idx[c] = mix(gid * kChains + c) % entries;
  uint64_t acc = 0;
  for (uint32_t h = 0; h < hops; ++h) {
    uint64_t e[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) e[c] = buf[idx[c]];
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
      acc += e[c];
      idx[c] = (e[c] ^ mix(static_cast<uint32_t>(e[c] >> 32) + h)) % entries;
    }
  }
  if (threadIdx.x == 0) lds[0] = static_cast<uint32_t>(acc);
  if (acc == 0x1234567ull) sink[gid] = acc;  // keeps the loads live
}

__global__ void fill_kernel(uint64_t* buf, uint64_t entries) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < entries; i += gridDim.x * 256ull)
    buf[i] = (static_cast<uint64_t>(mix(static_cast<uint32_t>(i))) << 32) |
             mix(static_cast<uint32_t>(i) ^ 0x9e3779b9u);
}

template <int kChains>
double run(const uint64_t* buf, uint64_t entries, uint32_t hops, int waves_per_simd, int cus,
           int reps, uint64_t* sink, size_t lds_per_cu) {
  // 256-thread workgroups (4 waves, one per SIMD): W workgroups per CU give
  // W waves per SIMD; the LDS reservation caps the residency at W
  const int wg_per_cu = waves_per_simd;
  const size_t lds = lds_per_cu / static_cast<size_t>(wg_per_cu) - 1024;
  const uint32_t blocks = static_cast<uint32_t>(cus * wg_per_cu);
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  chase_kernel<kChains><<<blocks, 256, lds>>>(buf, entries, hops, sink);
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  float best = 1e30f;
  for (int r = 0; r < reps; ++r) {
    CHECK(hipEventRecord(a));
    chase_kernel<kChains><<<blocks, 256, lds>>>(buf, entries, hops, sink);
    CHECK(hipEventRecord(b));
    CHECK(hipEventSynchronize(b));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, a, b));
    best = ms < best ? ms : best;
  }
  CHECK(hipEventDestroy(a));
  CHECK(hipEventDestroy(b));
  const double loads = static_cast<double>(blocks) * 256.0 * kChains * hops;
  return loads / (best / 1e3);
}

}  // namespace

int main(int argc, char** argv) {
  double mib = 160;
  uint32_t hops = 64;
  int reps = 3;
  for (int i = 1; i + 1 < argc; i += 2) {
    if (!std::strcmp(argv[i], "--mib")) mib = std::atof(argv[i + 1]);
    else if (!std::strcmp(argv[i], "--hops")) hops = static_cast<uint32_t>(std::atoi(argv[i + 1]));
    else if (!std::strcmp(argv[i], "--reps")) reps = std::atoi(argv[i + 1]);
  }
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  const int cus = prop.multiProcessorCount;
  const size_t lds_per_cu = 160 * 1024;
  const uint64_t entries = static_cast<uint64_t>(mib * 1024 * 1024 / 8);
  uint64_t* buf = nullptr;
  uint64_t* sink = nullptr;
  CHECK(hipMalloc(&buf, entries * 8));
  CHECK(hipMalloc(&sink, static_cast<size_t>(cus) * 16 * 256 * 8));
  fill_kernel<<<8192, 256>>>(buf, entries);
  CHECK(hipDeviceSynchronize());
  std::printf("buffer %.0f MiB (%llu entries), %u hops per chain, %d CUs\n", mib,
              static_cast<unsigned long long>(entries), hops, cus);
  std::printf("waves/SIMD chains      loads/s    GB/s@8B\n");
  for (int w : {1, 2, 4, 8}) {
    const double r1 = run<1>(buf, entries, hops, w, cus, reps, sink, lds_per_cu);
    const double r2 = run<2>(buf, entries, hops, w, cus, reps, sink, lds_per_cu);
    const double r4 = run<4>(buf, entries, hops, w, cus, reps, sink, lds_per_cu);
    std::printf("%10d %6d %12.3e %10.1f\n", w, 1, r1, r1 * 8 / 1e9);
    std::printf("%10d %6d %12.3e %10.1f\n", w, 2, r2, r2 * 8 / 1e9);
    std::printf("%10d %6d %12.3e %10.1f\n", w, 4, r4, r4 * 8 / 1e9);
    std::fflush(stdout);
  }
  CHECK(hipFree(buf));
  CHECK(hipFree(sink));
  return 0;
}
