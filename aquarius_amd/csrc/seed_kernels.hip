// seed_kernels.hip -- particle seeding on the GPU (gfx950), bit-identical to the host
// seeding of flip_seed.cpp.
//
// Fluid::createParticlesAt / createSquareFluid / createSphereFluid
// (Fluid/Fluid_3D_FLIP.cu:508-594): per visited cell a 2x2x2 lattice of jittered
// sub-particles, clamped to the domain and filtered by the volume. The host visits
// cells in a fixed loop order and draws three SplitMix64 values per sub-particle
// before the filter, so candidate c of a volume (cell k in loop order, sub o) uses
// draws base + 3c .. base + 3c + 2 of a counter-based stream: draw m is
// mix(seed + (m + 1) * golden). Every candidate is therefore independent; a count
// pass and a write pass with a stable in-block prefix keep the host's order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "aqflip_internal.h"

#pragma clang fp contract(off)

namespace aqf {

constexpr int kSeedBlock = 256;

__device__ __forceinline__ float draw01(uint64_t seed, uint64_t m) {
  uint64_t z = seed + (m + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (float)(z >> 40) * (1.0f / 16777216.0f);
}

// Candidate c of volume v: its position q and whether the volume keeps it.
__device__ __forceinline__ bool seed_candidate(const SeedVolume& v, const SeedParams& sp, long long c,
                                               float q[3]) {
  const long long k = c >> 3;
  const int o = (int)(c & 7);
  const int i2 = (int)(k % v.n[2]);
  const long long t = k / v.n[2];
  const int i1 = (int)(t % v.n[1]), i0 = (int)(t / v.n[1]);
  // loop order -> cell coordinates (box: z, x, y; sphere: y, x, z)
  int cell[3];
  cell[v.axis[0]] = v.o[0] + i0;
  cell[v.axis[1]] = v.o[1] + i1;
  cell[v.axis[2]] = v.o[2] + i2;
  const float spacing = sp.h / 2;
  const float half = spacing / 2;
  const float lattice[3] = {(o & 4) ? 0.5f : -0.5f, (o & 2) ? 0.5f : -0.5f, (o & 1) ? 0.5f : -0.5f};
  const uint64_t m0 = v.base + 3ull * (uint64_t)c;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float centre = ((float)cell[a] + 0.5f) * sp.h;
    const float jit = draw01(sp.seed, m0 + a) - 0.5f;
    const float sub = lattice[a] * spacing + centre;
    float val = sub + (jit * spacing) * sp.jitter;
    const float upper = (float)sp.S[a] * sp.h - half;
    val = fminf(upper, val);
    q[a] = fmaxf(half, val);
  }
  if (!v.sphere) return true;
  const float d0 = q[0] - v.cc[0], d1 = q[1] - v.cc[1], d2 = q[2] - v.cc[2];
  return __builtin_sqrtf(d0 * d0 + d1 * d1 + d2 * d2) < v.limit;
}

__device__ int seed_block_scan(int v, int* total) {
  __shared__ int wsum[kSeedBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kSeedBlock / 64; ++w) {
    if (w < wid) base += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

__global__ __launch_bounds__(kSeedBlock) void k_seed_count(SeedVolume v, SeedParams sp, long long nc,
                                                           int* __restrict__ counts) {
  const long long c = (long long)blockIdx.x * kSeedBlock + threadIdx.x;
  float q[3];
  const bool keep = c < nc && seed_candidate(v, sp, c, q);
  int tot;
  seed_block_scan(keep ? 1 : 0, &tot);
  if (threadIdx.x == 0) counts[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kSeedBlock) void k_seed_write(SeedVolume v, SeedParams sp, long long nc,
                                                           const int* __restrict__ offsets, long long out0,
                                                           Particles P) {
  const long long c = (long long)blockIdx.x * kSeedBlock + threadIdx.x;
  float q[3];
  const bool keep = c < nc && seed_candidate(v, sp, c, q);
  int tot;
  const int r = seed_block_scan(keep ? 1 : 0, &tot);
  if (!keep) return;
  const long long i = out0 + offsets[blockIdx.x] + r;
  float f[4] = {0.f, 0.f, 0.f, 0.f};
  if (v.phase >= 0 && v.phase < 4) f[v.phase] = 1.f;
  P.a[i] = make_float4(q[0], q[1], q[2], 0.f);
  P.b[i] = make_float4(0.f, 0.f, f[0], f[1]);
  P.c[i] = make_float4(f[2], f[3], 0.f, 0.f);
}

void launch_seed_count(const SeedVolume& v, const SeedParams& sp, long long nc, int* counts, hipStream_t s) {
  const long long nb = (nc + kSeedBlock - 1) / kSeedBlock;
  hipLaunchKernelGGL(k_seed_count, dim3((unsigned)nb), dim3(kSeedBlock), 0, s, v, sp, nc, counts);
}
void launch_seed_write(const SeedVolume& v, const SeedParams& sp, long long nc, const int* offsets,
                       long long out0, Particles P, hipStream_t s) {
  const long long nb = (nc + kSeedBlock - 1) / kSeedBlock;
  hipLaunchKernelGGL(k_seed_write, dim3((unsigned)nb), dim3(kSeedBlock), 0, s, v, sp, nc, offsets, out0, P);
}
int seed_block_size() { return kSeedBlock; }

}  // namespace aqf
