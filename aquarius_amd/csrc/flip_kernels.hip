// flip_kernels.hip -- CDNA4 (gfx950) kernels of the FLIP simulation step.
//
// Every kernel keeps the floating-point operation order of the reference kernel
// it replaces (cited per kernel), compiled without FMA contraction, so results
// match the CPU oracle bit for bit where the arithmetic allows. Iterative solvers
// read the previous iterate and write a new buffer (ping-pong) instead of the
// reference's racy in-place texture/surface update.
#include <hip/hip_runtime.h>

#include <climits>

#include "aqflip_internal.h"

#pragma clang fp contract(off)

namespace aqf {

constexpr int kBlock = 256;
constexpr int kXCD = 8;

// Workgroups of a launch are dispatched round-robin over the 8 XCDs. Map the
// hardware block id so that each XCD walks one contiguous range of logical
// blocks: a slab of cells (or a run of sorted particles) then stays in one XCD's
// L2 for the whole launch and across consecutive launches of the same shape.
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, bid = blockIdx.x;
  const int per = nb / kXCD, rem = nb % kXCD;
  const int xcd = bid % kXCD, k = bid / kXCD;
  return xcd * per + min(xcd, rem) + k;
}
__device__ __forceinline__ int gtid() { return xcd_block() * kBlock + (int)threadIdx.x; }
// xcd_block() for a launch of nb blocks passed as an argument (no hidden-argument load).
__device__ __forceinline__ int xcd_block_n(int nb) {
  const int per = nb / kXCD, rem = nb % kXCD, bid = blockIdx.x;
  const int xcd = bid % kXCD;
  return xcd * per + min(xcd, rem) + bid / kXCD;
}

static inline int nblocks(long long n) { return (int)((n + kBlock - 1) / kBlock); }

// ------------------------------------------------------------------ helpers
// trilinearUnitKernel (Fluid/WeightKernels.cuh:131-136): |r| > 1 ? 0 : 1 - |r|.
// Written as max(0, 1 - |r|), two instructions instead of three: for |r| <= 1 the
// two agree, and for |r| > 1 the difference 1 - |r| is negative (exactly, since
// |r| > 1 and 1 are both floats) so the max gives +0, the reference's 0. (They
// differ only for a NaN r, which no finite particle position produces.)
// max(0, 1 - |r|) written as a clamp to [0, 1] (1 - |r| <= 1; NaN gives 0 either way,
// with the DX10 clamp mode HIP kernels run in), which the compiler folds into the
// subtraction's clamp bit: one instruction instead of two.
__device__ __forceinline__ float hat_unit(float r) { return __builtin_amdgcn_fmed3f(1.f - fabsf(r), 0.f, 1.f); }
// Correctly rounded float division by a constant d from inv = 1/d rounded to
// double: the product x*inv carries a relative error below 2^-52, while the exact
// quotient of two floats is never a float rounding midpoint and lies more than
// 2^-49 (relative) away from every midpoint, so rounding the double product to
// float gives exactly the IEEE quotient x/d for every float x. Three instructions
// in place of the ~12 of the IEEE division sequence.
__device__ __forceinline__ float div_exact(float x, double inv) { return (float)((double)x * inv); }
// trilinearUnitKernel(r / h) with the exact division.
__device__ __forceinline__ float hat_div(float r, double inv_h) { return hat_unit(div_exact(r, inv_h)); }
__device__ __forceinline__ float lerpf_(float a, float b, float t) { return a + t * (b - a); }
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// Write-through store. A kernel's end writes every dirty line of each XCD's L2 back to
// memory, so that the next kernel, on any XCD, reads what it stored, and a short kernel
// waits for that write-back. A value stored at device scope (the sc1 bit: a relaxed
// agent-scope atomic store is a plain vector store with it) is written through the L2 as
// it is stored and leaves nothing to write back. Used by the pressure and diffusion
// sweeps, a few microseconds each, whose outputs the next sweep reads at once: 3.44 ->
// 3.27 us per pressure sweep at 128^3. Kernels that store more (P2G, the sort, the
// extrapolation, gravity, projection) were measured slower with it and store plainly.
template <typename T>
__device__ __forceinline__ void st_wt(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Interior cell (x,y,z) of a flat z-fastest index over X*Y*Z.
__device__ __forceinline__ void cell_of(const Geom& G, int n, int& x, int& y, int& z) {
  x = (int)fast_div((uint32_t)n, G.divYZ);
  const int r = n - x * (G.Y * G.Z);
  y = (int)fast_div((uint32_t)r, G.divZ);
  z = r - y * G.Z;
}
// Face cell (x,y,z) of a flat z-fastest index over (X+1)*(Y+1)*(Z+1).
__device__ __forceinline__ void face_of(const Geom& G, int n, int& x, int& y, int& z) {
  x = (int)fast_div((uint32_t)n, G.divYZ1);
  const int r = n - x * ((G.Y + 1) * (G.Z + 1));
  y = (int)fast_div((uint32_t)r, G.divZ1);
  z = r - y * (G.Z + 1);
}

// ------------------------------------------------------------------ fill
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}
void launch_fill_u32(void* p, uint32_t v, long long n, hipStream_t s) {
  int nb = (int)std::min<long long>((n + kBlock - 1) / kBlock, 8192);
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_fill_u32, dim3(nb), dim3(kBlock), 0, s, (uint32_t*)p, v, n);
}

// ------------------------------------------------------------------ hashing
// calcHashImpl (Fluid/Fluid_3D_kernels.cuh:8-27): truncating float->int of pos/dx.
// Positions are always inside [dx/4, L - dx/4] after advection and validated on
// upload, so the clamp to the grid is a no-op there; it only keeps a non-finite
// position (from a non-finite velocity) from indexing outside the bins.
__device__ __forceinline__ int cell_hash(const Geom& G, float px, float py, float pz) {
  const int x = clampi((int)div_exact(px, G.inv_h), 0, G.X - 1);
  const int y = clampi((int)div_exact(py, G.inv_h), 0, G.Y - 1);
  const int z = clampi((int)div_exact(pz, G.inv_h), 0, G.Z - 1);
  return x * (G.Y * G.Z) + y * G.Z + z;
}

// Bin one particle per lane: lanes of the wave holding the same hash reserve their
// slots with one atomic (issued by the lowest such lane) and take consecutive
// ranks in lane order. rank = slot of the particle inside its cell's segment
// before the segment sort. The groups are found first (ballots and shuffles only),
// then every group leader issues its atomic in the same instruction, so a wave
// waits for one atomic round trip however many cells its particles span.
// A group whose reservation does not start the cell's segment (base > 0) marks the cell:
// only cells with two or more runs can be out of order.
__device__ __forceinline__ void bin_particle(int hh, bool valid, int* __restrict__ count,
                                             uint8_t* __restrict__ multi, int& rank) {
  const int lane = (int)(threadIdx.x & 63);
  const unsigned long long lt = (1ull << lane) - 1ull;
  unsigned long long pending = __ballot(valid);
  int leader = lane;
  unsigned long long group = 0;
  while (pending) {
    const int l = __ffsll((long long)pending) - 1;
    const int lh = __shfl(hh, l, 64);
    const bool mine = valid && hh == lh;
    const unsigned long long same = __ballot(mine) & pending;
    if (mine) {
      leader = l;
      group = same;
    }
    pending &= ~same;
  }
  int base = 0;
  if (valid && lane == leader) {
    base = atomicAdd(&count[hh], __popcll(group));
    if (base > 0) multi[hh] = 1;
  }
  base = __shfl(base, leader, 64);
  rank = base + __popcll(group & lt);
}

__global__ void k_hash_bin(Geom G, Particles P, long long Np, int* __restrict__ hash,
                           int* __restrict__ rank, int* __restrict__ count, uint8_t* __restrict__ multi) {
  const int i = gtid();
  const bool valid = i < Np;
  int hh = 0;
  if (valid) {
    const float4 a = P.a[i];
    hh = cell_hash(G, a.x, a.y, a.z);
  }
  int r;
  bin_particle(hh, valid, count, multi, r);
  if (valid) {
    hash[i] = hh;
    rank[i] = r;
  }
}
void launch_hash_count(const Geom& G, Particles P, long long Np, Bins B, hipStream_t s) {
  hipLaunchKernelGGL(k_hash_bin, dim3(nblocks(Np)), dim3(kBlock), 0, s, G, P, Np, B.hash, B.rank,
                     B.count, B.multi);
}

// ------------------------------------------------------------------ scan
// Exclusive prefix sum of the per-cell counts: offsets[c] = first sorted slot of
// cell c, offsets[X*Y*Z] = Np. Three passes: chunk sums, scan of chunk sums in one
// block, chunk-local scan plus chunk base.
constexpr int kScanPerThread = 16;
constexpr int kScanChunk = kBlock * kScanPerThread;  // 4096 cells per block

__device__ int block_exclusive_scan(int v, int* total) {
  __shared__ int wsum[kBlock / 64];
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
  for (int w = 0; w < kBlock / 64; ++w) {
    if (w < wid) base += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

// A thread's kScanPerThread counts: 16-byte loads when all are in range (every
// thread but the last one's), else one predicated load each.
__device__ __forceinline__ void load_counts(const int* __restrict__ count, int n, int base,
                                            int v[kScanPerThread]) {
  if (base + kScanPerThread <= n) {
    const int4* __restrict__ q = reinterpret_cast<const int4*>(count + base);
#pragma unroll
    for (int k = 0; k < kScanPerThread / 4; ++k) {
      const int4 t = q[k];
      v[4 * k] = t.x, v[4 * k + 1] = t.y, v[4 * k + 2] = t.z, v[4 * k + 3] = t.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k) v[k] = base + k < n ? count[base + k] : 0;
  }
}
__global__ void k_scan_reduce(const int* __restrict__ count, int n, int* __restrict__ partials) {
  const int b = blockIdx.x;
  const int base = b * kScanChunk + threadIdx.x * kScanPerThread;
  int v[kScanPerThread];
  load_counts(count, n, base, v);
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) s += v[k];
  int tot;
  block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) partials[b] = tot;
}

// Exclusive scan of the block totals in place, by one block walking them 256 at a time
// (grids with more than kScanDirectParts blocks, where every final block re-summing the
// totals before it would grow with the square of the block count).
constexpr int kScanDirectParts = 1024;
__global__ void k_scan_parts(int* __restrict__ partials, int nparts) {
  int carry = 0;
  for (int j0 = 0; j0 < nparts; j0 += kBlock) {
    const int j = j0 + (int)threadIdx.x;
    const int v = j < nparts ? partials[j] : 0;
    int tot;
    const int ex = block_exclusive_scan(v, &tot);
    if (j < nparts) partials[j] = carry + ex;
    carry += tot;
  }
}

// Block b adds the totals of blocks 0..b-1 (read here, with its counts, in the same
// round trip; or, when `prefixed`, their sum already formed by k_scan_parts) to its own
// exclusive scan; the last block also writes offsets[n]. Integer sums, so any
// association gives the same offsets.
__global__ void k_scan_final(int* __restrict__ count, int n, const int* __restrict__ partials,
                             int nparts, int* __restrict__ offsets, int prefixed) {
  __shared__ int wsum[kBlock / 64], wpre[kBlock / 64];
  const int b = blockIdx.x;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int base = b * kScanChunk + threadIdx.x * kScanPerThread;
  int v[kScanPerThread];
  load_counts(count, n, base, v);
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) s += v[k];
  int pre = 0;
  if (!prefixed)
    for (int j = threadIdx.x; j < b; j += kBlock) pre += partials[j];
  else if (threadIdx.x == 0)
    pre = partials[b];  // the other threads add 0
  int x = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
    pre += __shfl_xor(pre, o, 64);
  }
  if (lane == 63) wsum[wid] = x;
  if (lane == 0) wpre[wid] = pre;
  __syncthreads();
  int below = 0, tot = 0, prefix = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    if (w < wid) below += wsum[w];
    tot += wsum[w];
    prefix += wpre[w];
  }
  int run = prefix + below + x - s;
  int o[kScanPerThread];
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    o[k] = run;
    run += v[k];
  }
  if (base + kScanPerThread <= n) {
    int4* __restrict__ q = reinterpret_cast<int4*>(offsets + base);
#pragma unroll
    for (int k = 0; k < kScanPerThread / 4; ++k)
      q[k] = make_int4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
  } else {
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k)
      if (base + k < n) offsets[base + k] = o[k];
  }
  if (b == nparts - 1 && threadIdx.x == 0) offsets[n] = prefix + tot;
  // The counts' last reader: the nonzero ones (the cells holding particles, a quarter of
  // the grid in the 128^3 scenes) are set back to zero, so the next binning needs no
  // clearing pass. (Zeroing all of them here cost as much as the clearing pass.)
  if (base + kScanPerThread <= n) {
    int4* __restrict__ q = reinterpret_cast<int4*>(count + base);
#pragma unroll
    for (int k = 0; k < kScanPerThread / 4; ++k)
      if ((v[4 * k] | v[4 * k + 1] | v[4 * k + 2] | v[4 * k + 3]) != 0) q[k] = make_int4(0, 0, 0, 0);
  } else {
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k)
      if (base + k < n && v[k] != 0) count[base + k] = 0;
  }
}

void launch_scan(const Geom& G, Bins B, hipStream_t s) {
  const int n = G.ncells_hash;
  const int nparts = (n + kScanChunk - 1) / kScanChunk;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nparts), dim3(kBlock), 0, s, B.count, n, B.partials);
  const int prefixed = nparts > kScanDirectParts ? 1 : 0;
  if (prefixed) hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(kBlock), 0, s, B.partials, nparts);
  hipLaunchKernelGGL(k_scan_final, dim3(nparts), dim3(kBlock), 0, s, B.count, n, B.partials, nparts,
                     B.offsets, prefixed);
}

// ------------------------------------------------------------------ stable counting sort
// sort_by_key(hash, index) (Fluid/Fluid_3D_common.cuh:35), stable by previous index.
// Each particle lands at offsets[hash] + rank; ranks are consecutive in index order
// within a wave, but the waves that share a cell reserve their runs in arbitrary
// order, so each segment is then sorted by previous index (usually already in
// order). applySortImpl (Fluid_3D_kernels.cuh:95-102) becomes a gather of the
// particle records through the permutation.
__global__ void k_scatter(const int* __restrict__ hash, const int* __restrict__ rank, long long Np,
                          const int* __restrict__ offsets, int* __restrict__ perm) {
  const int i = gtid();
  if (i >= Np) return;
  perm[offsets[hash[i]] + rank[i]] = i;
}

// Only cells that received two or more runs of ranks (flagged by bin_particle) can be
// out of order; every other thread reads its flag and returns, and every flag is
// cleared for the next binning. Segments of up to kSegFast entries, and of up to kSegMid,
// are loaded into registers in one round trip (predicated loads, INT_MAX past the end),
// sorted there by a bitonic network (every index a compile-time constant, so the array
// stays in registers) and written back once; the entries are distinct particle indices,
// so any correct sort gives the same order. Longer segments are checked by a loop and,
// when unsorted, insertion-sorted in place. (After the ball of the 128^3 multiphase
// scene meets the pool, 100-180 cells hold 17-30 particles: by the insertion sort, one
// dependent load and store per step, they took the kernel from 12.5 to 81 us.)
constexpr int kSegFast = 16;
constexpr int kSegMid = 32;
template <int N>
__device__ __forceinline__ void bitonic_sort(int (&v)[N]) {
#pragma unroll
  for (int k = 2; k <= N; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int l = i ^ j;
        if (l > i) {
          const int lo = min(v[i], v[l]), hi = max(v[i], v[l]);
          const bool up = (i & k) == 0;
          v[i] = up ? lo : hi;
          v[l] = up ? hi : lo;
        }
      }
}
// Segment [s, e) with e - s <= N, sorted in registers unless already in order.
template <int N>
__device__ __forceinline__ void sort_segment(int* __restrict__ perm, int s, int e) {
  int v[N];
#pragma unroll
  for (int k = 0; k < N; ++k) v[k] = s + k < e ? perm[s + k] : INT_MAX;
  bool sorted = true;
#pragma unroll
  for (int k = 1; k < N; ++k) sorted &= s + k >= e || v[k - 1] < v[k];
  if (sorted) return;
  bitonic_sort<N>(v);
#pragma unroll
  for (int k = 0; k < N; ++k)
    if (s + k < e) perm[s + k] = v[k];
}
__global__ void k_segsort(const int* __restrict__ offsets, int ncells, int* __restrict__ perm,
                          uint8_t* __restrict__ multi) {
  const int c = gtid();
  if (c >= ncells || !multi[c]) return;
  multi[c] = 0;
  const int s = offsets[c], e = offsets[c + 1];
  if (e - s <= kSegFast) {
    sort_segment<kSegFast>(perm, s, e);
    return;
  }
  if (e - s <= kSegMid) {
    sort_segment<kSegMid>(perm, s, e);
    return;
  }
  bool sorted = true;
  for (int i = s + 1; i < e; ++i) sorted &= perm[i - 1] < perm[i];
  if (sorted) return;
  for (int i = s + 1; i < e; ++i) {
    const int v = perm[i];
    int j = i - 1;
    while (j >= s && perm[j] > v) {
      perm[j + 1] = perm[j];
      --j;
    }
    perm[j + 1] = v;
  }
}

template <int K>
__global__ void k_gather(Particles cur, Particles alt, long long Np, const int* __restrict__ perm) {
  const int k = gtid();
  if (k >= Np) return;
  const int j = perm[k];
  alt.a[k] = cur.a[j];
  alt.b[k] = cur.b[j];
  if (K > 2) alt.c[k] = cur.c[j];
}

void launch_scatter_sort_gather(const Geom& G, Particles cur, Particles alt, long long Np, int K,
                                Bins B, hipStream_t s) {
  hipLaunchKernelGGL(k_scatter, dim3(nblocks(Np)), dim3(kBlock), 0, s, B.hash, B.rank, Np,
                     B.offsets, B.perm);
  hipLaunchKernelGGL(k_segsort, dim3(nblocks(G.ncells_hash)), dim3(kBlock), 0, s, B.offsets,
                     G.ncells_hash, B.perm, B.multi);
  if (K > 2)
    hipLaunchKernelGGL(k_gather<4>, dim3(nblocks(Np)), dim3(kBlock), 0, s, cur, alt, Np, B.perm);
  else
    hipLaunchKernelGGL(k_gather<2>, dim3(nblocks(Np)), dim3(kBlock), 0, s, cur, alt, Np, B.perm);
}

// ------------------------------------------------------------------ standalone spatial hashing
// performSpatialHashing2 (Fluid/Fluid_3D_common.cuh:28-56) for any particle layout:
// positions are the first three floats of records of `stride` floats.
__global__ void k_hash_records(Geom G, const float* __restrict__ pos, long long n, int stride,
                               int* __restrict__ hash, int* __restrict__ rank,
                               int* __restrict__ count, uint8_t* __restrict__ multi) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  int hh = 0;
  if (valid) {
    const float* p = pos + i * stride;
    hh = cell_hash(G, p[0], p[1], p[2]);
  }
  int r;
  bin_particle(hh, valid, count, multi, r);
  if (valid) {
    hash[i] = hh;
    rank[i] = r;
  }
}
void launch_hash_records(const Geom& G, const float* pos, long long n, int stride, Bins B, hipStream_t s) {
  hipLaunchKernelGGL(k_hash_records, dim3(nblocks(n)), dim3(kBlock), 0, s, G, pos, n, stride, B.hash,
                     B.rank, B.count, B.multi);
}
void launch_scatter_segsort(const Geom& G, long long n, Bins B, hipStream_t s) {
  hipLaunchKernelGGL(k_scatter, dim3(nblocks(n)), dim3(kBlock), 0, s, B.hash, B.rank, n, B.offsets,
                     B.perm);
  hipLaunchKernelGGL(k_segsort, dim3(nblocks(G.ncells_hash)), dim3(kBlock), 0, s, B.offsets,
                     G.ncells_hash, B.perm, B.multi);
}
// findCellStartEndImpl (Fluid/Fluid_kernels.cuh:3-22) from the offsets: first and
// last (inclusive) sorted slot of every cell, -1 for empty cells.
__global__ void k_cell_ranges(int ncells, const int* __restrict__ offsets, int* __restrict__ start,
                              int* __restrict__ end) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int s = offsets[c], e = offsets[c + 1];
  if (start) start[c] = e > s ? s : -1;
  if (end) end[c] = e > s ? e - 1 : -1;
}
// Sorted hashes and, optionally, the records in sorted order (applySortImpl,
// Fluid/Fluid_3D_kernels.cuh:95-102).
__global__ void k_sorted_outputs(long long n, const int* __restrict__ perm, const int* __restrict__ hash,
                                 int* __restrict__ sorted_hash, const float* __restrict__ rec,
                                 float* __restrict__ sorted_rec, int stride) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int j = perm[k];
  if (sorted_hash) sorted_hash[k] = hash[j];
  if (sorted_rec)
    for (int f = 0; f < stride; ++f) sorted_rec[k * stride + f] = rec[(long long)j * stride + f];
}
void launch_hash_outputs(const Geom& G, long long n, int stride, Bins B, const float* rec, int* start, int* end,
                         int* sorted_hash, float* sorted_rec, hipStream_t s) {
  if (start || end)
    hipLaunchKernelGGL(k_cell_ranges, dim3(nblocks(G.ncells_hash)), dim3(kBlock), 0, s, G.ncells_hash,
                       B.offsets, start, end);
  if ((sorted_hash || sorted_rec) && n > 0)
    hipLaunchKernelGGL(k_sorted_outputs, dim3(nblocks(n)), dim3(kBlock), 0, s, n, B.perm, B.hash,
                       sorted_hash, rec, sorted_rec, stride);
}

// ------------------------------------------------------------------ P2G
// Fuses resetCellImpl (Fluid_3D_FLIP.cu:287-302), transferToCellAccumPhase (:8-69),
// transferToCellDividePhase (:72-116) and transferVolumeFractionsToCell (:118-165).
//
// One wavefront per block owns a segment of kP2GCells cells along z of one (x, y)
// row, two lanes per cell. The reference visits, for each cell, its 27 neighbours in
// (dx, dy, dz) order by flattened index h = (x+dx)*Y*Z + (y+dy)*Z + z+dz
// (range-checked in 1D only) and each neighbour's particles in sorted order. For a
// fixed (dx, dy) the three cells dz = -1, 0, 1 are consecutive hashes, so their
// particles form one contiguous sorted range per cell, and the ranges of the
// segment's cells form one contiguous range per wave (row wrap-around included,
// exactly as the flattened index wraps). The wave stages that range through LDS with
// coalesced 16-byte loads, 256 particles per chunk (about the whole window of 28
// cells, so every lane has work in every chunk); because every lane shares x and y,
// the x and y hat factors of a staged particle are the same for the whole wave and
// are multiplied once, as (kx * ky), the first product of trilinearHatKernel. Each
// lane then forms its z factor and accumulates its cell's range in order, so every
// sum sees the reference's operands in the reference's order. Outside the centre slab
// (dx, dy) = (0, 0) about half of the staged particles have x/y factor products that
// are all exactly zero; they only add +-0 to sums that start at +0, which changes no
// bit for finite velocities, so staging drops them (wave ballot, order kept) and each
// lane maps its range onto the kept particles.
//
// Lane 2i (role A) accumulates the u and v faces of cell i (weights with kzC) and the
// cell-centre fractions; lane 2i + 1 (role B) the w face (weights with kzF). Both run
// the same instructions on their own LDS records, A = (pu, pv, pz, vx) and
// B = (pw, vy, pz, vz), both taking the second velocity from B.y = vy (role B's second
// sum is meaningless and dropped). The centre slab stages every particle, so the
// own-cell fractions are read from the particle record in memory. (One lane per cell
// over 64-cell segments left half the lanes idle in each chunk: 441 against 391 us.)
//
// Latency: the bin offsets of all nine slabs arrive in one round trip per wave (not
// one per slab), and the records of the next chunk are loaded into registers while
// the current one is accumulated, so those loads overlap the accumulation. The kernel
// is then bound by instruction issue; the hat is one clamped subtraction and the
// accumulation loop takes four slots per trip.
constexpr int kP2GWave = 64;
constexpr int kP2GChunk = 256;  // particles staged per pass
constexpr int kP2GCells = 28;   // cells per wave

struct P2GGeom {
  int nsegz;
  FastDiv divSeg, divY;
};

template <int K>
__global__ __launch_bounds__(kP2GWave) void k_p2g(Geom G, P2GGeom PG, Particles P,
                                                   const int* __restrict__ off, Grid gr, int nv,
                                                   int hv, int nf, StepParams sp) {
  // Role records; the first array padded by four records (64 bytes), so a cell's two
  // lanes, which read slot c of each, hit banks 16 apart instead of the same banks.
  __shared__ float4 sR[2][kP2GChunk + 4];
  // Kept records before each staged position (0..kP2GChunk): a lane's window maps onto
  // the kept records with two lookups.
  __shared__ uint16_t sPre[kP2GChunk + 1];
  const int b = xcd_block();
  const int xy = (int)fast_div((uint32_t)b, PG.divSeg);
  const int zs = b - xy * PG.nsegz;
  const int x = (int)fast_div((uint32_t)xy, PG.divY);
  const int y = xy - x * G.Y;
  const int lane = threadIdx.x;
  const int ci = lane >> 1, role = lane & 1;
  const int z0 = zs * kP2GCells;
  const int z = z0 + ci;
  const int zl = min(z0 + kP2GCells, G.Z) - 1;  // last cell of the segment
  const bool active = ci < kP2GCells && z < G.Z;
  const float h = G.h;
  const double ih = G.inv_h;
  const float fx = (float)x, fy = (float)y, fz = (float)z;
  const float xF = fx * h, yF = fy * h, zF = fz * h;
  const float xC = (fx + 0.5f) * h, yC = (fy + 0.5f) * h, zC = (fz + 0.5f) * h;
  const float zsel = role ? zF : zC;
  const float4* __restrict__ myR = sR[role];
  const int NC = G.ncells_hash;
  const int YZ = G.Y * G.Z;
  float t1 = 0.f, t2 = 0.f, s1 = 0.f, s2 = 0.f;
  float f[K];
#pragma unroll
  for (int a = 0; a < K; ++a) f[a] = 0.f;
  float ws = 0.f;
  // Every bin offset the wave needs, in one round trip: slab s = 3 (dx + 1) + dy + 1
  // holds off[clamp(h0 + z0 - 1 + j)], j = 0..31, in lanes 32 (s & 1) + j of wb[s >> 1]
  // (h0 = (x + dx) Y Z + (y + dy) Z). Lane j = ci gives the cell's window start, ci + 3
  // its end, 0 and zl - z0 + 3 the wave's; slab 4 at ci + 1, ci + 2 is the own cell.
  int wb[5];
#pragma unroll
  for (int r = 0; r < 5; ++r) {
    const int sl = 2 * r + (lane >> 5);
    wb[r] = 0;
    if (sl < 9) {
      const int h0 = (x + sl / 3 - 1) * YZ + (y + sl % 3 - 1) * G.Z;
      wb[r] = off[clampi(h0 + z0 - 1 + (lane & 31), 0, NC)];
    }
  }
  // (shuffles run on every lane: a lane outside the segment may be the source)
  const int cj = min(ci, kP2GCells - 1);
  int ownLo = __shfl(wb[2], cj + 1, 64), ownHi = __shfl(wb[2], cj + 2, 64);
  if (!active) ownLo = ownHi = 0;
  const int wend = zl - z0 + 3;
  auto slab_offsets = [&](int sl) {  // wave-uniform sl
    const int r = sl >> 1;
    return r == 0 ? wb[0] : r == 1 ? wb[1] : r == 2 ? wb[2] : r == 3 ? wb[3] : wb[4];
  };
  auto win_lo = [&](int sl) { return __builtin_amdgcn_readlane(slab_offsets(sl), (sl & 1) << 5); };
  auto win_hi = [&](int sl) { return __builtin_amdgcn_readlane(slab_offsets(sl), ((sl & 1) << 5) + wend); };
  uint32_t busy = 1u << 9;  // slabs with particles in the window; bit 9 ends the walk
#pragma unroll
  for (int sl = 0; sl < 9; ++sl)
    if (win_lo(sl) < win_hi(sl)) busy |= 1u << sl;
  // The wave walks (slab, chunk) items in the reference's slab order. The records of the
  // next item are loaded into registers while the current chunk is accumulated from LDS
  // (after it, for the centre slab, whose accumulation reads fractions from memory).
  float4 LA[kP2GChunk / kP2GWave];
  float2 LB[kP2GChunk / kP2GWave];
  // (unconditional loads, clamped to the item's last particle, so none waits in a branch)
  auto load_item = [&](int sl, int cs) {
    const int last = min(cs + kP2GChunk, win_hi(sl)) - 1;
#pragma unroll
    for (int bi = 0; bi < kP2GChunk / kP2GWave; ++bi) {
      const int i = min(cs + bi * kP2GWave + lane, last);
      LA[bi] = P.a[i];
      LB[bi] = reinterpret_cast<const float2*>(P.b)[2 * i];  // (vy, vz)
    }
  };
  int sl = __builtin_ctz(busy);
  int cs = sl < 9 ? win_lo(sl) : 0;
  if (sl < 9) load_item(sl, cs);
  while (sl < 9) {
    const int base = (sl & 1) << 5;
    const int v = slab_offsets(sl);
    const int whi = __builtin_amdgcn_readlane(v, base + wend);
    int lo = __shfl(v, base + cj, 64), hi = __shfl(v, base + cj + 3, 64);
    if (!active) lo = hi = 0;
    const bool centre = sl == 4;  // (dx, dy) = (0, 0)
    const int ce = min(cs + kP2GChunk, whi);
    const int nst = ce - cs;
    int nsl = sl, ncs = ce;
    if (ce >= whi) {
      nsl = __builtin_ctz(busy & (~0u << (sl + 1)));
      ncs = nsl < 9 ? win_lo(nsl) : 0;
    }
    __syncthreads();
    int nkept = 0;
#pragma unroll
    for (int bi = 0; bi < kP2GChunk / kP2GWave; ++bi) {
      if (bi * kP2GWave >= nst) continue;
      const int k = bi * kP2GWave + lane;
      bool keep = false;
      const float4 A = LA[bi];
      const float2 B = LB[bi];
      float pu = 0.f, pv = 0.f, pw = 0.f;
      if (k < nst) {
        const float kxF = hat_div(A.x - xF, ih), kxC = hat_div(A.x - xC, ih);
        const float kyF = hat_div(A.y - yF, ih), kyC = hat_div(A.y - yC, ih);
        pu = kxF * kyC;
        pv = kxC * kyF;
        pw = kxC * kyC;
        keep = centre || pu != 0.f || pv != 0.f || pw != 0.f;
      }
      const uint64_t m = __ballot(keep);
      const int q = nkept + __popcll(m & ((1ull << lane) - 1));  // kept before position k
      sPre[k] = (uint16_t)q;
      if (keep) {
        // Narrow stores of the loaded registers as they are: a 16-byte store of values
        // from two loads would copy them first, and a copy waits for the load.
        float* r0 = reinterpret_cast<float*>(&sR[0][q]);
        float* r1 = reinterpret_cast<float*>(&sR[1][q]);
        *reinterpret_cast<float2*>(r0) = make_float2(pu, pv);
        asm volatile("" ::: "memory");
        *reinterpret_cast<float2*>(r0 + 2) = make_float2(A.z, A.w);
        asm volatile("" ::: "memory");
        r1[0] = pw;
        asm volatile("" ::: "memory");
        r1[1] = B.x;
        asm volatile("" ::: "memory");
        r1[2] = A.z;
        asm volatile("" ::: "memory");
        r1[3] = B.y;
      }
      nkept += __popcll(m);
    }
    if (lane == 0) sPre[nst] = (uint16_t)nkept;  // nst a multiple of 64: no lane wrote it
    if (!centre && nsl < 9) load_item(nsl, ncs);
    __syncthreads();
    const int c0 = sPre[clampi(lo - cs, 0, nst)];
    const int c1 = sPre[clampi(hi - cs, 0, nst)];
    auto accumulate = [&](int c, bool ownc) {
      const float4 R = myR[c];
      const float vy = sR[1][c].y;
      const float kz = hat_div(R.z - zsel, ih);
      const float w1 = R.x * kz;  // A: u face (x, y+.5, z+.5); B: w face (x+.5, y+.5, z)
      const float w2 = R.y * kz;  // A: v face (x+.5, y, z+.5); B: unused
      t1 += w1 * R.w;
      t2 += w2 * vy;
      s1 += w1;
      s2 += w2;
      if (ownc) {  // role A: cell centre (x+.5, y+.5, z+.5), weight pw * kzC
        const float wc = sR[1][c].x * kz;
        const float4 Bq = P.b[cs + c];  // centre slab: slot c is particle cs + c
        float fr[4] = {Bq.z, Bq.w, 0.f, 0.f};
        if (K > 2) {
          const float4 Cq = P.c[cs + c];
          fr[2] = Cq.x;
          fr[3] = Cq.y;
        }
#pragma unroll
        for (int a = 0; a < K; ++a) f[a] += wc * fr[a];
        ws += wc;
      }
    };
    // Slots [ca, cb) in order, four per trip (the count mod 4 first).
    auto accumulate_range = [&](int ca, int cb) {
      int c = ca;
      if ((cb - ca) & 1) accumulate(c++, false);
      if ((cb - ca) & 2) {
        accumulate(c, false);
        accumulate(c + 1, false);
        c += 2;
      }
      for (; c < cb; c += 4) {
        accumulate(c, false);
        accumulate(c + 1, false);
        accumulate(c + 2, false);
        accumulate(c + 3, false);
      }
    };
    if (!centre) {
      accumulate_range(c0, c1);
    } else {
      const int o0 = clampi(ownLo - cs, c0, c1), o1 = clampi(ownHi - cs, o0, c1);
      accumulate_range(c0, o0);
      for (int c = o0; c < o1; ++c) accumulate(c, true);
      accumulate_range(o1, c1);
      if (nsl < 9) load_item(nsl, ncs);
    }
    sl = nsl;
    cs = ncs;
  }
  // role B's w-face results to its role-A partner
  const float tw = __shfl_xor(t1, 1, 64), sw = __shfl_xor(s1, 1, 64);
  // A segment holding FLUID stamps the face rows within 4 rows in x and in y of its row:
  // the only rows the step's last two extrapolation passes can change (k_extrapolate).
  if (__ballot(active && role == 0 && ownHi > ownLo)) {
    const uint32_t stamp = *gr.stepstamp;
#pragma unroll
    for (int k0 = 0; k0 < 81; k0 += kP2GWave) {
      const int k = k0 + lane;
      if (k < 81) {
        const int fx = clampi(x + k / 9 - 4, 0, G.X), fy = clampi(y + k % 9 - 4, 0, G.Y);
        gr.rowstamp[fx * (G.Y + 1) + fy] = stamp;
      }
    }
  }
  if (!active || role) return;
  float tv0 = t1, tv1 = t2, tv2 = tw;
  uint8_t hvb = 0;
  if (s1 > 0.f) { tv0 /= s1; hvb |= 1; }
  if (s2 > 0.f) { tv1 /= s2; hvb |= 2; }
  if (sw > 0.f) { tv2 /= sw; hvb |= 4; }
  const long long gi = G.g(x, y, z);
  gr.vel[0][gi] = tv0;
  gr.vel[1][gi] = tv1;
  gr.vel[2][gi] = tv2;
  gr.nvel[nv][0][gi] = tv0;
  gr.nvel[nv][1][gi] = tv1;
  gr.nvel[nv][2][gi] = tv2;
  gr.hasvel[hv][gi] = hvb;
  gr.content[gi] = (ownHi > ownLo) ? FLUID : AIR;
  if (ws > 0.f) {
#pragma unroll
    for (int a = 0; a < K; ++a) f[a] /= ws;
    float sum = f[0];
#pragma unroll
    for (int a = 1; a < K; ++a) sum += f[a];
#pragma unroll
    for (int a = 0; a < K; ++a) f[a] /= sum;
  }
  float dens = sp.rho[0] * f[0];
#pragma unroll
  for (int a = 1; a < K; ++a) dens += sp.rho[a] * f[a];
  gr.density[gi] = dens;
#pragma unroll
  for (int a = 0; a < K; ++a) {
    gr.vf[a][gi] = f[a];
    gr.nvf[nf][a][gi] = f[a];
  }
}

void launch_p2g(const Geom& G, Particles P, long long Np, int K, Bins B, Grid gr, int nv, int hv,
                int nf, StepParams sp, hipStream_t s) {
  (void)Np;
  P2GGeom PG;
  PG.nsegz = (G.Z + kP2GCells - 1) / kP2GCells;
  PG.divSeg = make_fastdiv((uint32_t)PG.nsegz);
  PG.divY = make_fastdiv((uint32_t)G.Y);
  const int nb = G.X * G.Y * PG.nsegz;
  switch (K) {
    case 1: hipLaunchKernelGGL(k_p2g<1>, dim3(nb), dim3(kP2GWave), 0, s, G, PG, P, B.offsets, gr, nv, hv, nf, sp); break;
    case 2: hipLaunchKernelGGL(k_p2g<2>, dim3(nb), dim3(kP2GWave), 0, s, G, PG, P, B.offsets, gr, nv, hv, nf, sp); break;
    case 3: hipLaunchKernelGGL(k_p2g<3>, dim3(nb), dim3(kP2GWave), 0, s, G, PG, P, B.offsets, gr, nv, hv, nf, sp); break;
    default: hipLaunchKernelGGL(k_p2g<4>, dim3(nb), dim3(kP2GWave), 0, s, G, PG, P, B.offsets, gr, nv, hv, nf, sp); break;
  }
}

// ------------------------------------------------------------------ gravity + boundary
// applyGravityImpl (Fluid/Fluid_3D_kernels.cu:4-39) gathered per target face cell,
// then fixBoundaryX/Y/Z (Fluid_3D_kernels.cu:44-98) with the reference's pair
// coverage (Fluid_3D_common.cu:13-41). Runs over the whole (X+1)^3 face extent,
// since the padding layer is both a gravity target and a boundary face.
// Every load of a face (its components and content, and the content of its -x, -y, -z
// neighbours, read at the face itself where out of range) is issued before any is used.
__global__ void k_gravity_boundary(float* __restrict__ U, float* __restrict__ V, float* __restrict__ W,
                                   uint8_t* __restrict__ content, int doGravity, int doBoundary, int nb,
                                   StepParams sp, Geom G) {
  asm volatile("" ::"s"(G.X), "s"(G.Y), "s"(G.Z), "s"(G.EY), "s"(G.EZ), "s"(G.sx), "s"(G.nfaces),
               "s"(G.divZ1.m), "s"(G.divZ1.s), "s"(G.divYZ1.m), "s"(G.divYZ1.s), "s"(sp.dt), "s"(sp.g[0]),
               "s"(sp.g[1]), "s"(sp.g[2]), "s"(sp.coverX), "s"(sp.coverY), "s"(sp.coverZ));
  const int n = xcd_block_n(nb) * kBlock + (int)threadIdx.x;
  if (n >= G.nfaces) return;
  int x, y, z;
  face_of(G, n, x, y, z);
  const long long gi = G.g(x, y, z);
  const uint8_t ct = content[gi];
  const float u0 = U[gi], v0 = V[gi], w0 = W[gi];
  float u = u0, v = v0, w = w0;
  long long ox = x >= 1 ? G.sx : 0, oy = y >= 1 ? (long long)G.EZ : 0, oz = z >= 1 ? 1 : 0;
  asm volatile("" : "+v"(ox), "+v"(oy), "+v"(oz));  // unconditional loads, one round trip
  const uint8_t cxm = content[gi - ox], cym = content[gi - oy], czm = content[gi - oz];
  asm volatile("" ::"v"(u), "v"(v), "v"(w), "v"((uint32_t)ct), "v"((uint32_t)cxm), "v"((uint32_t)cym),
               "v"((uint32_t)czm));
  if (doGravity) {
    const float ax = sp.g[0] * sp.dt, ay = sp.g[1] * sp.dt, az = sp.g[2] * sp.dt;
    const bool inter = x < G.X && y < G.Y && z < G.Z;
    if (inter && ct == FLUID) {
      u += ax;
      v += ay;
      w += az;
    }
    if (ct == AIR) {
      if (sp.g[0] != 0.f && x >= 1 && x - 1 < G.X && y < G.Y && z < G.Z && cxm == FLUID) {
        u += ax;
        v += ay;
        w += az;
      }
      if (sp.g[1] != 0.f && y >= 1 && x < G.X && y - 1 < G.Y && z < G.Z && cym == FLUID) {
        u += ax;
        v += ay;
        w += az;
      }
      if (sp.g[2] != 0.f && z >= 1 && x < G.X && y < G.Y && z - 1 < G.Z && czm == FLUID) {
        u += ax;
        v += ay;
        w += az;
      }
    }
  }
  uint8_t nc = ct;
  if (doBoundary) {
    const bool cx = (y * (G.Z + 1) + z) < sp.coverX;
    const bool cy = (x * (G.Z + 1) + z) < sp.coverY;
    const bool cz = (x * (G.Y + 1) + y) < sp.coverZ;
    if (cx && x == 0) u = fmaxf(0.f, u);
    if (cx && x == G.X) { u = fminf(0.f, u); nc = SOLID; }
    if (cy && y == 0) v = fmaxf(0.f, v);
    if (cy && y == G.Y) { v = fminf(0.f, v); nc = SOLID; }
    if (cz && z == 0) w = fmaxf(0.f, w);
    if (cz && z == G.Z) { w = fminf(0.f, w); nc = SOLID; }
  }
  // (only what changed is stored: most faces, away from FLUID and the walls, keep their
  // values, and the same bits need no store)
  if (__float_as_uint(u) != __float_as_uint(u0)) U[gi] = u;
  if (__float_as_uint(v) != __float_as_uint(v0)) V[gi] = v;
  if (__float_as_uint(w) != __float_as_uint(w0)) W[gi] = w;
  if (nc != ct) content[gi] = nc;
}
void launch_gravity_boundary(const Geom& G, Grid gr, int nv, StepParams sp, int doGravity,
                             int doBoundary, hipStream_t s) {
  const int nb = nblocks(G.nfaces);
  hipLaunchKernelGGL(k_gravity_boundary, dim3(nb), dim3(kBlock), 0, s, gr.nvel[nv][0], gr.nvel[nv][1],
                     gr.nvel[nv][2], gr.content, doGravity, doBoundary, nb, sp, G);
}

// ------------------------------------------------------------------ divergence
// countParticlesInCellsImpl (Fluid_3D_FLIP.cu:168-185): cellEnd - cellStart with an
// inclusive end, i.e. count-1 for occupied cells and 0 for empty ones; then
// computeDivergenceImpl (Fluid_3D_kernels.cu:426-458) with restPPC = 8, whose
// correction term is evaluated in double as in the reference. Also stores the
// Jacobi right-hand side -div*density (Fluid_3D_kernels.cu:474-479).
// Marks the launch block of the calling thread active if any of its cells is FLUID.
// Every thread of the block must reach it (no early return before).
__device__ __forceinline__ void mark_active(const Grid& gr, bool fluid) {
  const int any = __syncthreads_or(fluid ? 1 : 0);
  if (threadIdx.x == 0) gr.active[xcd_block()] = any ? 1 : 0;
}

// Every load of a cell (content, the six face components, its two bin offsets, density)
// is issued before the block-wide FLUID test: one memory round trip per cell.
struct DivergenceFields {
  const float* U;
  const float* V;
  const float* W;
  const float* density;
  float* div;
  float* rhs;
  uint32_t* active;
};
__global__ void k_divergence(const uint8_t* __restrict__ content, const int* __restrict__ off, int nb,
                             DivergenceFields f, Geom G) {
  asm volatile("" ::"s"(G.X), "s"(G.Y), "s"(G.Z), "s"(G.EY), "s"(G.EZ), "s"(G.sx), "s"(G.ncells_hash),
               "s"(G.divZ.m), "s"(G.divZ.s), "s"(G.divYZ.m), "s"(G.divYZ.s), "s"(f.U), "s"(f.V), "s"(f.W),
               "s"(f.density), "s"(f.div), "s"(f.rhs), "s"(f.active));
  const int blk = xcd_block_n(nb);
  const int n = blk * kBlock + (int)threadIdx.x;
  const bool in = n < G.ncells_hash;
  const int nn = in ? n : 0;  // past the end: cell 0's loads, unused
  int x, y, z;
  cell_of(G, nn, x, y, z);
  const long long gi = G.g(x, y, z);
  const float* __restrict__ U = f.U;
  const float* __restrict__ V = f.V;
  const float* __restrict__ W = f.W;
  const uint8_t c = content[gi];
  const float v1 = V[gi + G.EZ], v0 = V[gi], u1 = U[gi + G.sx], u0 = U[gi], w1 = W[gi + 1], w0 = W[gi];
  const int o0 = off[nn], o1 = off[nn + 1];
  const float dens = f.density[gi];
  asm volatile("" ::"v"((uint32_t)c), "v"(v1), "v"(v0), "v"(u1), "v"(u0), "v"(w1), "v"(w0), "v"(o0), "v"(o1),
               "v"(dens));
  const int any = __syncthreads_or(in && c == FLUID ? 1 : 0);  // the block's active flag
  if (threadIdx.x == 0) f.active[blk] = any ? 1 : 0;
  if (!in) return;
  float div = v1 - v0 + u1 - u0 + w1 - w0;
  const int cnt = o1 - o0;
  const int pc = cnt > 0 ? cnt - 1 : 0;
  const double corr = (double)((float)pc - 8.0f) * 0.001;
  div = (float)((double)div - corr);
  f.div[gi] = div;
  const bool air = c == AIR;
  f.rhs[gi] = air ? __uint_as_float(kAirRhsBits) : -div * dens;
}

// Right-hand side from divergence, density and content (after a field upload).
__global__ void k_refresh_rhs(Geom G, Grid gr) {
  const int n = gtid();
  const bool in = n < G.ncells_hash;
  int fx = 0, fy = 0, fz = 0;
  if (in) cell_of(G, n, fx, fy, fz);
  mark_active(gr, in && gr.content[G.g(fx, fy, fz)] == FLUID);
  if (!in) return;
  int x, y, z;
  cell_of(G, n, x, y, z);
  const long long gi = G.g(x, y, z);
  gr.rhs[gi] = gr.content[gi] == AIR ? __uint_as_float(kAirRhsBits) : -gr.div[gi] * gr.density[gi];
}
void launch_refresh_rhs(const Geom& G, Grid gr, hipStream_t s) {
  hipLaunchKernelGGL(k_refresh_rhs, dim3(nblocks(G.ncells_hash)), dim3(kBlock), 0, s, G, gr);
  launch_active_list(G, gr, s);
}
void launch_divergence(const Geom& G, Grid gr, int nv, Bins B, hipStream_t s) {
  DivergenceFields f;
  f.U = gr.nvel[nv][0];
  f.V = gr.nvel[nv][1];
  f.W = gr.nvel[nv][2];
  f.density = gr.density;
  f.div = gr.div;
  f.rhs = gr.rhs;
  f.active = gr.active;
  const int nb = nblocks(G.ncells_hash);
  hipLaunchKernelGGL(k_divergence, dim3(nb), dim3(kBlock), 0, s, gr.content, B.offsets, nb, f, G);
  launch_active_list(G, gr, s);
}

// ------------------------------------------------------------------ Jacobi
__global__ void k_fill_interior_zero(Geom G, float* __restrict__ p) {
  const int n = gtid();
  if (n >= G.ncells_hash) return;
  int x, y, z;
  cell_of(G, n, x, y, z);
  p[G.g(x, y, z)] = 0.f;
}
void launch_fill_pressure_zero(const Geom& G, Grid gr, int pb, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_interior_zero, dim3(nblocks(G.ncells_hash)), dim3(kBlock), 0, s, G,
                     gr.p[pb]);
}

// Cells per thread of the pressure sweep kernels: a thread issues the loads of all
// its cells before it uses any, and a launch needs 1/kSweepCells of the
// workgroups. At 128^3 two cells per thread took a pressure sweep from 5.30 to
// 4.83 us and a paired sweep from 8.75 to 8.31 us; four were slower (6.1 us).
constexpr int kSweepCells = 2;
constexpr int kSweepSpan = kSweepCells * kBlock;  // cells per chunk (one workgroup's work)
// Workgroups of a listed sweep (multiples of kXCD): they stride over the active list.
// Measured at 128^3 (1,441 of 4,096 chunks active): a pressure sweep is fastest with
// 1024 (3.59 us; 2048: 3.79); a paired sweep that computes its diffusion part takes
// 6.19 us with 1536-2048 workgroups, 6.24 with 1024 and 7.28 with 512, while one that
// skips it runs the pressure sweep alone, 4.56 us with 1024 against 4.66 with 2048
// (graph replay of the 128^3 multiphase workload: 794 against 786 steps/s). With the
// sweeps' write-through stores, 768 / 1024 / 1280 workgroups: pressure sweep 3.52 / 3.29
// / 3.33 us, 863 / 901 / 898 steps/s.
constexpr int kListedGrid = 1024;
constexpr int kListedGridPaired = 1024;

int sweep_chunks(const Geom& G) { return (G.ncells_hash + kSweepSpan - 1) / kSweepSpan; }

// A chunk holds FLUID if any of the 256-cell blocks it spans does (flags of
// k_divergence / k_refresh_rhs; the slot past the last block is always 0).
__device__ __forceinline__ bool chunk_active(const uint32_t* __restrict__ active, int ch) {
  bool any = false;
#pragma unroll
  for (int c = 0; c < kSweepCells; ++c) any = any || active[ch * kSweepCells + c] != 0;
  return any;
}

// The active chunks in ascending order: list[0] = count, list[1..count] = chunk
// ids. One 1024-thread workgroup: each thread takes a contiguous range of chunks and
// reads their flags into registers in batches of kListBatch (one round trip per
// batch), then a wave-shuffle scan of the per-thread counts places its ids.
// Rebuilt after every update of the flags.
constexpr int kListThreads = 1024;
constexpr int kListBatch = 8;
// Work split of a listed sweep over `total` list entries on `grid` workgroups: XCD k
// (workgroups k, k + groups, ...) takes entries [k per, (k + 1) per), its workgroups
// striding through them by nj.
struct ListedSplit {
  int groups, nj, per;
};
__host__ __device__ inline ListedSplit listed_split(int grid, int total) {
  const int groups = grid >= kXCD ? kXCD : 1;
  return ListedSplit{groups, grid / groups, (total + groups - 1) / groups};
}
// Bit k: chunk b + k (< c1) holds FLUID. A chunk's kSweepCells = 2 flags are one
// 8-byte load, unconditional (index clamped to the thread's last chunk) so the batch's
// loads are in flight together.
static_assert(kSweepCells == 2, "one 8-byte load per chunk's flags");
__device__ __forceinline__ uint32_t batch_mask(const uint32_t* __restrict__ active, int b, int c1) {
  uint2 f[kListBatch];
#pragma unroll
  for (int k = 0; k < kListBatch; ++k) f[k] = *reinterpret_cast<const uint2*>(active + 2 * min(b + k, c1 - 1));
  uint32_t m = 0;
#pragma unroll
  for (int k = 0; k < kListBatch; ++k)
    if (b + k < c1 && (f[k].x | f[k].y) != 0u) m |= 1u << k;
  return m;
}
__global__ __launch_bounds__(kListThreads) void k_active_list(int nchunks, const uint32_t* __restrict__ active,
                                                              uint32_t* __restrict__ list,
                                                              uint32_t* __restrict__ dchange,
                                                              uint4* __restrict__ btab, int gridJ,
                                                              int gridP) {
  __shared__ int wsum[kListThreads / 64];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int per = (nchunks + kListThreads - 1) / kListThreads;
  const int c0 = min(t * per, nchunks), c1 = min(c0 + per, nchunks);
  int n = 0;
  uint32_t first = 0;  // the first batch's mask, kept for the second pass
  for (int b = c0; b < c1; b += kListBatch) {
    const uint32_t m = batch_mask(active, b, c1);
    if (b == c0) first = m;
    n += __popc(m);
  }
  int x = n;  // inclusive scan within the wave, then across the waves
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  int base = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kListThreads / 64; ++w) {
    base += w < wid ? wsum[w] : 0;
    total += wsum[w];
  }
  int pos = base + x - n;
  for (int b = c0; b < c1; b += kListBatch) {
    const uint32_t m = b == c0 ? first : batch_mask(active, b, c1);
#pragma unroll
    for (int k = 0; k < kListBatch; ++k)
      if (m & (1u << k)) list[1 + pos++] = (uint32_t)(b + k);
  }
  if (t == 0) list[0] = (uint32_t)total;
  // The sweeps' per-workgroup table: {chunks, first list position, first two ids},
  // read back from the list this workgroup has just written.
  __syncthreads();
  const ListedSplit LJ = listed_split(gridJ, total), LP = listed_split(gridP, total);
  for (int e = t; e < gridJ + gridP; e += kListThreads) {
    const bool pj = e < gridJ;
    const ListedSplit& L = pj ? LJ : LP;
    const int b = pj ? e : e - gridJ;
    const int g = b % L.groups, j = b / L.groups;
    const int t0 = g * L.per + j, end = min((g + 1) * L.per, total);
    const int cnt = t0 < end ? (end - t0 + L.nj - 1) / L.nj : 0;
    const uint32_t id0 = cnt > 0 ? list[1 + t0] : 0u, id1 = cnt > 1 ? list[1 + t0 + L.nj] : 0u;
    btab[e] = make_uint4((uint32_t)cnt, (uint32_t)t0, id0, id1);
  }
  if (t == 0) dchange[0] = 0u;  // the first paired sweep of the next step records here
}
int sweep_grid(const Geom& G, bool full, int listed);
void launch_active_list(const Geom& G, Grid gr, hipStream_t s) {
  hipLaunchKernelGGL(k_active_list, dim3(1), dim3(kListThreads), 0, s, sweep_chunks(G), gr.active, gr.alist,
                     gr.dchange, gr.btab, gr.listed[0], gr.listed[1]);
}
int sweep_table_entries(const Geom& G) {
  return sweep_grid(G, false, kListedGrid) + sweep_grid(G, false, kListedGridPaired);
}
void set_listed_grids(const Geom& G, Grid& gr, int limit) {
  const int lj = limit > 0 ? std::max(kXCD, std::min(limit, kListedGrid)) : kListedGrid;
  const int lp = limit > 0 ? std::max(kXCD, std::min(limit, kListedGridPaired)) : kListedGridPaired;
  gr.listed[0] = sweep_grid(G, false, lj);
  gr.listed[1] = sweep_grid(G, false, lp);
}

// Work split of a sweep. A full sweep (every cell written) runs one workgroup per
// chunk. A listed sweep visits only the active chunks: XCD k (workgroups
// k, k + kXCD, ..., the hardware's round-robin placement) takes the k-th eighth of
// the list, its workgroups striding through it, so a sweep costs workgroups only
// where FLUID is and each XCD's L2 sees one contiguous range of the grid. At 128^3
// (35 % of the chunks active) this took a pressure sweep from 4.88 to 3.62 us.
int sweep_grid(const Geom& G, bool full, int listed = kListedGrid) {
  const int nch = sweep_chunks(G);
  if (full || nch < kXCD) return nch;
  return std::min(nch, listed) / kXCD * kXCD;
}
// Cell c (0..kSweepCells-1) of this thread in chunk ch, its grid index and its six
// clamp-to-edge neighbours in the reference's order x+1, x-1, y+1, y-1, z+1, z-1.
// Threads past the last cell get cell 0 (in bounds) and ok = false.
struct SweepCell {
  long long gi, nb[6];
  bool ok;
};
__device__ __forceinline__ SweepCell sweep_cell(const Geom& G, int ch, int c) {
  SweepCell q;
  const int n = ch * kSweepSpan + c * kBlock + (int)threadIdx.x;
  q.ok = n < G.ncells_hash;
  int x, y, z;
  cell_of(G, q.ok ? n : 0, x, y, z);
  q.gi = G.g(x, y, z);
  q.nb[0] = q.gi + (x < G.X - 1 ? G.sx : 0);
  q.nb[1] = q.gi - (x > 0 ? G.sx : 0);
  q.nb[2] = q.gi + (y < G.Y - 1 ? (long long)G.EZ : 0);
  q.nb[3] = q.gi - (y > 0 ? (long long)G.EZ : 0);
  q.nb[4] = q.gi + (z < G.Z - 1 ? 1 : 0);
  q.nb[5] = q.gi - (z > 0 ? 1 : 0);
  return q;
}
// jacobiImpl (Fluid/Fluid_3D_kernels.cu:460-505) for one cell whose loads are done:
// AIR cells 0, otherwise (x+1 + x-1 + y+1 + y-1 + z+1 + z-1 + rhs) / 6 summed in
// that order. AIR cells are recognised from the right-hand side's reserved bit
// pattern.
__device__ __forceinline__ float jacobi_value(float b, const float v[6]) {
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 6; ++k) s += v[k];
  s += b;
  return __float_as_uint(b) == kAirRhsBits ? 0.f : div_exact(s, 1.0 / 6.0);
}

// One Jacobi sweep, read-old/write-new. All loads of a thread's cells are issued
// before any is used (an empty `asm volatile` consumes them, so the compiler cannot
// sink loads behind a branch), so a cell costs one memory round trip. AIR cells are
// always written (0): what a full sweep needs and what a listed sweep's buffer
// already holds; a listed sweep visits only the chunks with FLUID. With `pzero` the
// input is the zero guess: the neighbours are 0 and are not read (same sums).
template <bool PZERO>
__device__ __forceinline__ void jacobi_chunk(const Geom& G, const float* __restrict__ rhs,
                                             const float* __restrict__ pin, float* __restrict__ pout,
                                             int ch) {
  SweepCell q[kSweepCells];
  float b[kSweepCells], v[kSweepCells][6];
#pragma unroll
  for (int c = 0; c < kSweepCells; ++c) {
    q[c] = sweep_cell(G, ch, c);
    b[c] = rhs[q[c].gi];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[c][k] = PZERO ? 0.f : pin[q[c].nb[k]];
  }
#pragma unroll
  for (int c = 0; c < kSweepCells; ++c)
    asm volatile("" ::"v"(b[c]), "v"(v[c][0]), "v"(v[c][1]), "v"(v[c][2]), "v"(v[c][3]),
                 "v"(v[c][4]), "v"(v[c][5]));
#pragma unroll
  for (int c = 0; c < kSweepCells; ++c)
    if (q[c].ok) st_wt(&pout[q[c].gi], jacobi_value(b[c], v[c]));
}
// Values a sweep's prologue needs, in scalar registers at one point: all of their loads
// (kernel arguments, the workgroup's table entry) then issue before one wait, instead of
// one dependent round trip per group of uses. One asm statement, so that no register is
// reused (and its load ordered after an earlier use) among them.
#define AQF_SGPR_READY(G, e, ...)                                                               \
  asm volatile("" ::"s"(e.x), "s"(e.y), "s"(e.z), "s"(e.w), "s"(G.X), "s"(G.Y), "s"(G.Z),    \
               "s"(G.EY), "s"(G.EZ), "s"(G.sx), "s"(G.ncells_hash), "s"(G.divZ.m), "s"(G.divZ.s), \
               "s"(G.divYZ.m), "s"(G.divYZ.s), __VA_ARGS__)
template <bool PZERO>
__global__ void k_jacobi(const uint4* __restrict__ tab, const uint32_t* __restrict__ list,
                         const float* __restrict__ rhs, const float* __restrict__ pin,
                         float* __restrict__ pout, int full, int nb, Geom G) {
  const uint4 e = tab[full ? 0 : blockIdx.x];
  const int nj = nb >= kXCD ? nb / kXCD : nb;
  AQF_SGPR_READY(G, e, "s"(nj), "s"(nb));
  if (full) {
    const int per = nb / kXCD, rem = nb % kXCD, bid = blockIdx.x;
    const int xcd = bid % kXCD, k = bid / kXCD;
    jacobi_chunk<PZERO>(G, rhs, pin, pout, xcd * per + min(xcd, rem) + k);
    return;
  }
  const int n = (int)e.x;
  for (int k = 0; k < n; ++k)
    jacobi_chunk<PZERO>(G, rhs, pin, pout, (int)(k == 0 ? e.z : k == 1 ? e.w : list[1 + e.y + k * nj]));
}
void launch_jacobi(const Geom& G, Grid gr, int pin, int full, int pzero, hipStream_t s) {
  const dim3 grid(full || pzero ? sweep_grid(G, true) : gr.listed[0]);
  const int nb = (int)grid.x;
  if (pzero)
    hipLaunchKernelGGL(k_jacobi<true>, grid, dim3(kBlock), 0, s, gr.btab, gr.alist, gr.rhs, gr.p[pin],
                       gr.p[pin ^ 1], 1, nb, G);
  else
    hipLaunchKernelGGL(k_jacobi<false>, grid, dim3(kBlock), 0, s, gr.btab, gr.alist, gr.rhs, gr.p[pin],
                       gr.p[pin ^ 1], full, nb, G);
}

// diffusionJacobiImpl writes pressure = 0 on AIR cells (Fluid_3D_kernels.cu:529-532).
// Jacobi already leaves AIR cells at 0, so this only matters with 0 pressure sweeps.
__global__ void k_air_pressure_zero(Geom G, const uint8_t* __restrict__ content,
                                    float* __restrict__ p) {
  const int n = gtid();
  if (n >= G.ncells_hash) return;
  int x, y, z;
  cell_of(G, n, x, y, z);
  const long long gi = G.g(x, y, z);
  if (content[gi] == AIR) p[gi] = 0.f;
}
void launch_air_pressure_zero(const Geom& G, Grid gr, int pb, hipStream_t s) {
  hipLaunchKernelGGL(k_air_pressure_zero, dim3(nblocks(G.ncells_hash)), dim3(kBlock), 0, s, G,
                     gr.content, gr.p[pb]);
}

// ------------------------------------------------------------------ projection
// updateVelocityWithPressureImpl (Fluid/Fluid_3D_kernels.cu:118-207): faces with a
// FLUID side get u -= (p - p_other)/rho_used; hasVelocity is overwritten with
// "face touches fluid".
// Every load of a cell (its content, pressure and density, those of its -x, -y, -z
// neighbours, read at the cell itself where out of range, and its three face components)
// is issued before any is used: one memory round trip per cell.
struct ProjectFields {
  float* F[3];
  uint8_t* hv;
};
__global__ void k_project(const uint8_t* __restrict__ content, const float* __restrict__ p,
                          const float* __restrict__ density, int nb, ProjectFields f, Geom G) {
  asm volatile("" ::"s"(G.X), "s"(G.Y), "s"(G.Z), "s"(G.EY), "s"(G.EZ), "s"(G.sx), "s"(G.ncells_hash),
               "s"(G.divZ.m), "s"(G.divZ.s), "s"(G.divYZ.m), "s"(G.divYZ.s), "s"(f.F[0]), "s"(f.F[1]),
               "s"(f.F[2]), "s"(f.hv));
  const int n = xcd_block_n(nb) * kBlock + (int)threadIdx.x;
  if (n >= G.ncells_hash) return;
  int x, y, z;
  cell_of(G, n, x, y, z);
  const long long gi = G.g(x, y, z);
  const uint8_t tc = content[gi];
  const float tp = p[gi];
  const float td = density[gi];
  const int coord[3] = {x, y, z};
  const long long step[3] = {G.sx, (long long)G.EZ, 1};
  uint32_t oc[3];
  float op[3], od[3], fv[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    long long o = coord[d] > 0 ? step[d] : 0;
    asm volatile("" : "+v"(o));  // unconditional loads, one round trip
    const long long go = gi - o;
    oc[d] = content[go];
    op[d] = p[go];
    od[d] = density[go];
    fv[d] = f.F[d][gi];
  }
  asm volatile("" ::"v"((uint32_t)tc), "v"(tp), "v"(td));
#pragma unroll
  for (int d = 0; d < 3; ++d) asm volatile("" ::"v"(oc[d]), "v"(op[d]), "v"(od[d]), "v"(fv[d]));
  uint8_t hvb = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (coord[d] > 0) {
      if (tc == FLUID || oc[d] == FLUID) {
        float used = td;
        if (tc == FLUID && oc[d] == FLUID)
          used = (od[d] + td) / 2.f;
        else if (oc[d] == FLUID)
          used = od[d];
        f.F[d][gi] = fv[d] - (tp - op[d]) / used;
        hvb |= (uint8_t)(1u << d);
      }
    }
  }
  f.hv[gi] = hvb;
}
void launch_project(const Geom& G, Grid gr, int nv, int hv, int pb, hipStream_t s) {
  ProjectFields f;
  for (int d = 0; d < 3; ++d) f.F[d] = gr.nvel[nv][d];
  f.hv = gr.hasvel[hv];
  const int nb = nblocks(G.ncells_hash);
  hipLaunchKernelGGL(k_project, dim3(nb), dim3(kBlock), 0, s, gr.content, gr.p[pb], gr.density, nb, f, G);
}

// ------------------------------------------------------------------ extrapolation
// extrapolateVelocityByOne (Fluid/Fluid_3D_kernels.cu:210-422), read-old/write-new.
// A face without velocity takes the mean of the same component of its six
// neighbours that have velocity and whose velocity along the neighbour direction
// passes the reference's epsilon test (epsilon = -1 + 1e-6). Faces outside the
// launch domain are copied so both buffers stay complete.
//
// Inside a step the last two passes visit only the face rows P2G stamped (near_rows).
// That gives the same fields: projection rewrites every interior valid bit from FLUID
// adjacency alone, so after it only faces within one row (in x and in y) of a cell row
// holding FLUID have velocity, and each pass reaches one row further; rows beyond 4
// never change. The step's first pass runs over every face, which leaves those rows
// equal in both buffers, so the passes that skip them leave them as they would be.
// The pointers and integers come ahead of the geometry (preloaded into scalar registers),
// the grid size is an argument, and every scalar the prologue needs is named in one asm
// statement: a wave's argument and geometry loads, and the step stamp, arrive in one
// round trip, so a wave outside the near rows costs that and its row-stamp load.
struct ExtrapFields {
  const float* Ui;
  const float* Vi;
  const float* Wi;
  float* Uo;
  float* Vo;
  float* Wo;
  const uint8_t* Hi;
  uint8_t* Ho;
  const float* old[3];  // velocity, read with write_nvo
  float2* nvo[3];       // with write_nvo: (new value, velocity) of every visited face
};
// A face's row is its index divided by Z + 1 (x-major, z-fastest faces). With Z + 1 >= 64
// a wave spans at most two rows, and a wave whose rows both lack this step's stamp leaves
// after its first (scalar) round trip; `wide_rows` says the launcher found Z + 1 >= 64.
__global__ void k_extrapolate(const uint32_t* __restrict__ rowstamp, const uint32_t* __restrict__ stepstamp,
                              int near_rows, int nb, int nfaces, uint32_t z1m, uint32_t z1s, int wide_rows,
                              int write_nvo, ExtrapFields f, Geom G) {
  const int per = nb / kXCD, rem = nb % kXCD, bid = blockIdx.x;
  const int xcd = bid % kXCD;
  const int n = (xcd * per + min(xcd, rem) + bid / kXCD) * kBlock + (int)threadIdx.x;
  const FastDiv dz1{0u, z1m, z1s};
  uint32_t stamp = 0u, s0 = 0u, s1 = 0u;
  if (near_rows) {
    stamp = *stepstamp;
    if (wide_rows) {  // the wave's first and last rows, from scalar registers
      const int w0 = __builtin_amdgcn_readfirstlane(n - (int)(threadIdx.x & 63));
      s0 = rowstamp[fast_div((uint32_t)min(w0, nfaces - 1), dz1)];
      s1 = rowstamp[fast_div((uint32_t)min(w0 + 63, nfaces - 1), dz1)];
    }
  }
  asm volatile("" ::"s"(G.X), "s"(G.Y), "s"(G.Z), "s"(G.EY), "s"(G.EZ), "s"(G.sx), "s"(G.divYZ1.m),
               "s"(G.divYZ1.s), "s"(G.divZ1.m), "s"(G.divZ1.s), "s"(stamp), "s"(s0), "s"(s1), "s"(f.Ui),
               "s"(f.Vi), "s"(f.Wi), "s"(f.Uo), "s"(f.Vo), "s"(f.Wo), "s"(f.Hi), "s"(f.Ho));
  if (near_rows && wide_rows && s0 != stamp && s1 != stamp) return;
  if (n >= nfaces) return;
  int x, y, z;
  face_of(G, n, x, y, z);
  if (near_rows && rowstamp[fast_div((uint32_t)n, dz1)] != stamp) return;
  const long long gi = G.g(x, y, z);
  const float* __restrict__ Ui = f.Ui;
  const float* __restrict__ Vi = f.Vi;
  const float* __restrict__ Wi = f.Wi;
  float* __restrict__ Uo = f.Uo;
  float* __restrict__ Vo = f.Vo;
  float* __restrict__ Wo = f.Wo;
  const uint8_t* __restrict__ Hi = f.Hi;
  uint8_t* __restrict__ Ho = f.Ho;
  float ov[3] = {0.f, 0.f, 0.f};
  if (write_nvo) {
#pragma unroll
    for (int c = 0; c < 3; ++c) ov[c] = f.old[c][gi];
  }
  if (x >= G.X || y >= G.Y || z >= G.Z) {
    const float u = Ui[gi], v = Vi[gi], w = Wi[gi];
    asm volatile("" ::"v"(u), "v"(v), "v"(w), "v"(ov[0]), "v"(ov[1]), "v"(ov[2]));  // loads, then stores
    Uo[gi] = u;
    Vo[gi] = v;
    Wo[gi] = w;
    if (write_nvo) {
      f.nvo[0][gi] = make_float2(u, ov[0]);
      f.nvo[1][gi] = make_float2(v, ov[1]);
      f.nvo[2][gi] = make_float2(w, ov[2]);
    }
    return;
  }
  const float eps = (float)(-1 + 1e-6);
  const float* __restrict__ F[3] = {Ui, Vi, Wi};
  // neighbours -x, -y, -z, +x, +y, +z
  const bool ok[6] = {x > 0, y > 0, z > 0, x < G.X - 1, y < G.Y - 1, z < G.Z - 1};
  const long long off[6] = {-G.sx, -(long long)G.EZ, -1, G.sx, (long long)G.EZ, 1};
  // Every load (the face, and the valid bits and three components of each in-range
  // neighbour; out-of-range ones read the face itself and are ignored) is issued
  // before any is used: one memory round trip per face.
  const uint8_t H = Hi[gi];
  float val[3] = {Ui[gi], Vi[gi], Wi[gi]};
  uint32_t nh[6];
  float nv[6][3];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    // (the offset is hidden from the compiler, which would otherwise reuse the face's
    // own loads for an out-of-range neighbour: wait for them, then load the others)
    long long o = ok[k] ? off[k] : 0;
    asm volatile("" : "+v"(o));
    const long long go = gi + o;
    nh[k] = Hi[go];
#pragma unroll
    for (int c = 0; c < 3; ++c) nv[k][c] = F[c][go];
  }
  asm volatile("" ::"v"(val[0]), "v"(val[1]), "v"(val[2]), "v"(ov[0]), "v"(ov[1]), "v"(ov[2]));
#pragma unroll
  for (int k = 0; k < 6; ++k) asm volatile("" ::"v"(nh[k]), "v"(nv[k][0]), "v"(nv[k][1]), "v"(nv[k][2]));
  uint8_t Hn = H;
  if (H != 7) {
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
      if (H & (1u << comp)) continue;
      float sum = 0.f;
      int cnt = 0;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const float dirv = nv[k][k % 3];
        const bool dirOk = (k < 3) ? (dirv > eps) : (dirv < -eps);
        if (ok[k] && (nh[k] & (1u << comp)) && dirOk) {
          sum += nv[k][comp];
          cnt++;
        }
      }
      if (cnt > 0) {
        const double inv_cnt[7] = {0.0, 1.0, 1.0 / 2.0, 1.0 / 3.0, 1.0 / 4.0, 1.0 / 5.0, 1.0 / 6.0};
        val[comp] = div_exact(sum, inv_cnt[cnt]);
        Hn |= (uint8_t)(1u << comp);
      }
    }
  }
  Uo[gi] = val[0];
  Vo[gi] = val[1];
  Wo[gi] = val[2];
  Ho[gi] = Hn;
  if (write_nvo) {
#pragma unroll
    for (int c = 0; c < 3; ++c) f.nvo[c][gi] = make_float2(val[c], ov[c]);
  }
}
// gr.nvo in full: (newVelocity buffer nv, velocity) of every face.
__global__ void k_build_nvo(Geom G, const float* __restrict__ n0, const float* __restrict__ n1,
                            const float* __restrict__ n2, const float* __restrict__ o0,
                            const float* __restrict__ o1, const float* __restrict__ o2, float2* __restrict__ d0,
                            float2* __restrict__ d1, float2* __restrict__ d2) {
  const int n = gtid();
  if (n >= G.nfaces) return;
  int x, y, z;
  face_of(G, n, x, y, z);
  const long long gi = G.g(x, y, z);
  d0[gi] = make_float2(n0[gi], o0[gi]);
  d1[gi] = make_float2(n1[gi], o1[gi]);
  d2[gi] = make_float2(n2[gi], o2[gi]);
}
void launch_build_nvo(const Geom& G, Grid gr, int nv, hipStream_t s) {
  hipLaunchKernelGGL(k_build_nvo, dim3(nblocks(G.nfaces)), dim3(kBlock), 0, s, G, gr.nvel[nv][0], gr.nvel[nv][1],
                     gr.nvel[nv][2], gr.vel[0], gr.vel[1], gr.vel[2], gr.nvo[0], gr.nvo[1], gr.nvo[2]);
}
void launch_extrapolate(const Geom& G, Grid gr, int nvin, int hvin, bool near_rows, hipStream_t s,
                        bool write_nvo) {
  ExtrapFields f;
  f.Ui = gr.nvel[nvin][0];
  f.Vi = gr.nvel[nvin][1];
  f.Wi = gr.nvel[nvin][2];
  f.Uo = gr.nvel[nvin ^ 1][0];
  f.Vo = gr.nvel[nvin ^ 1][1];
  f.Wo = gr.nvel[nvin ^ 1][2];
  f.Hi = gr.hasvel[hvin];
  f.Ho = gr.hasvel[hvin ^ 1];
  for (int c = 0; c < 3; ++c) {
    f.old[c] = gr.vel[c];
    f.nvo[c] = gr.nvo[c];
  }
  const int nb = nblocks(G.nfaces);
  hipLaunchKernelGGL(k_extrapolate, dim3(nb), dim3(kBlock), 0, s, gr.rowstamp, gr.stepstamp, near_rows ? 1 : 0,
                     nb, G.nfaces, G.divZ1.m, G.divZ1.s, G.Z + 1 >= 64 ? 1 : 0, write_nvo ? 1 : 0, f, G);
}

// ------------------------------------------------------------------ diffusion
// diffusionJacobiImpl (Fluid/Fluid_3D_kernels.cu:520-556), read-old/write-new:
// f' = (lambda * sum of the six clamp-to-edge neighbours + f) / (1 + 2 lambda),
// then normalised by its channel sum; AIR cells keep their value.
template <int K>
__device__ __forceinline__ void diffusion_cell(const Geom& G, const Grid& gr, int fin, float lambda,
                                               double inv_cc, int n, int full) {
  int x, y, z;
  cell_of(G, n, x, y, z);
  const long long gi = G.g(x, y, z);
  const long long nb[6] = {gi + (x < G.X - 1 ? G.sx : 0), gi - (x > 0 ? G.sx : 0),
                           gi + (y < G.Y - 1 ? (long long)G.EZ : 0),
                           gi - (y > 0 ? (long long)G.EZ : 0), gi + (z < G.Z - 1 ? 1 : 0),
                           gi - (z > 0 ? 1 : 0)};
  // Every load (content, the cell's and its six neighbours' channels) is issued
  // before the AIR test: one memory round trip per cell.
  const uint8_t c = gr.content[gi];
  float v[7][K];
#pragma unroll
  for (int k = 0; k < 6; ++k)
#pragma unroll
    for (int a = 0; a < K; ++a) v[k][a] = gr.nvf[fin][a][nb[k]];
#pragma unroll
  for (int a = 0; a < K; ++a) v[6][a] = gr.nvf[fin][a][gi];
#pragma unroll
  for (int k = 0; k < 7; ++k)
#pragma unroll
    for (int a = 0; a < K; ++a) asm volatile("" ::"v"(v[k][a]));
  if (c == AIR) {
    if (full) {
#pragma unroll
      for (int a = 0; a < K; ++a) st_wt(&gr.nvf[fin ^ 1][a][gi], v[6][a]);
    }
    return;
  }
  float r[K];
#pragma unroll
  for (int a = 0; a < K; ++a) r[a] = 0.f;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
#pragma unroll
    for (int a = 0; a < K; ++a) r[a] += v[k][a] * lambda;
  }
#pragma unroll
  for (int a = 0; a < K; ++a) r[a] += v[6][a];
#pragma unroll
  for (int a = 0; a < K; ++a) r[a] = div_exact(r[a], inv_cc);  // r / (1 + 2 lambda)
  float sum = r[0];
#pragma unroll
  for (int a = 1; a < K; ++a) sum += r[a];
#pragma unroll
  for (int a = 0; a < K; ++a) st_wt(&gr.nvf[fin ^ 1][a][gi], r[a] / sum);
}

template <int K>
__global__ void k_diffusion(Geom G, Grid gr, int fin, float lambda, double inv_cc, int full) {
  if (!full && !gr.active[xcd_block()]) return;
  const int n = gtid();
  if (n >= G.ncells_hash) return;
  diffusion_cell<K>(G, gr, fin, lambda, inv_cc, n, full);
}
void launch_diffusion(const Geom& G, Grid gr, int K, int fin, float lambda, double inv_cc, int full,
                      hipStream_t s) {
  const int nb = nblocks(G.ncells_hash);
  switch (K) {
    case 1: hipLaunchKernelGGL(k_diffusion<1>, dim3(nb), dim3(kBlock), 0, s, G, gr, fin, lambda, inv_cc, full); break;
    case 2: hipLaunchKernelGGL(k_diffusion<2>, dim3(nb), dim3(kBlock), 0, s, G, gr, fin, lambda, inv_cc, full); break;
    case 3: hipLaunchKernelGGL(k_diffusion<3>, dim3(nb), dim3(kBlock), 0, s, G, gr, fin, lambda, inv_cc, full); break;
    default: hipLaunchKernelGGL(k_diffusion<4>, dim3(nb), dim3(kBlock), 0, s, G, gr, fin, lambda, inv_cc, full); break;
  }
}

// One Jacobi sweep and one diffusion sweep in a single launch. The two chains touch
// disjoint fields (pressure, and the volume fractions), so a thread runs both for
// its cells with the arithmetic of jacobi_value and diffusion_cell, computing the
// shared cell and neighbour indices once. Every load of both is issued before any
// is used, so the cells still cost one memory round trip. fullJ and fullD mean what
// `full` means for each sweep.
// The fields a paired sweep touches, passed as plain pointers: a kernel taking the whole
// Grid by value read its fields from the argument segment one dependent round trip at a
// time, after the branches that use them.
struct PairedFields {
  const float* rhs;
  const float* pin;
  float* pout;
  const uint8_t* content;
  const uint32_t* active;
  const float* fin[4];
  float* fout[4];
};
template <int K, bool PZERO>
__device__ __forceinline__ void paired_chunk(const Geom& G, const PairedFields& gr, int pin, int fin, float lambda,
                                             double inv_cc, bool fullD, bool doJ, bool doD, int ch,
                                             bool& changed) {
  (void)pin;
  (void)fin;
  const float* __restrict__ P = gr.pin;
  SweepCell q[kSweepCells];
  float b[kSweepCells], pv[kSweepCells][6], v[kSweepCells][7][K];
  uint8_t cc[kSweepCells];
#pragma unroll
  for (int c = 0; c < kSweepCells; ++c) {
    q[c] = sweep_cell(G, ch, c);
    b[c] = gr.rhs[q[c].gi];
#pragma unroll
    for (int k = 0; k < 6; ++k) pv[c][k] = PZERO ? 0.f : P[q[c].nb[k]];
    cc[c] = gr.content[q[c].gi];
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
      for (int a = 0; a < K; ++a) v[c][k][a] = gr.fin[a][q[c].nb[k]];
#pragma unroll
    for (int a = 0; a < K; ++a) v[c][6][a] = gr.fin[a][q[c].gi];
  }
#pragma unroll
  for (int c = 0; c < kSweepCells; ++c) {
    asm volatile("" ::"v"(b[c]), "v"(pv[c][0]), "v"(pv[c][1]), "v"(pv[c][2]), "v"(pv[c][3]),
                 "v"(pv[c][4]), "v"(pv[c][5]));
#pragma unroll
    for (int k = 0; k < 7; ++k)
#pragma unroll
      for (int a = 0; a < K; ++a) asm volatile("" ::"v"(v[c][k][a]));
  }
#pragma unroll
  for (int c = 0; c < kSweepCells; ++c) {
    if (!q[c].ok) continue;
    const long long gi = q[c].gi;
    if (doJ) st_wt(&gr.pout[gi], jacobi_value(b[c], pv[c]));
    if (!doD) continue;
    if (cc[c] == AIR) {  // diffusion_cell: AIR cells keep their value
      if (fullD) {
#pragma unroll
        for (int a = 0; a < K; ++a) st_wt(&gr.fout[a][gi], v[c][6][a]);
      }
      continue;
    }
    float r[K];
#pragma unroll
    for (int a = 0; a < K; ++a) r[a] = 0.f;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
#pragma unroll
      for (int a = 0; a < K; ++a) r[a] += v[c][k][a] * lambda;
    }
#pragma unroll
    for (int a = 0; a < K; ++a) r[a] += v[c][6][a];
#pragma unroll
    for (int a = 0; a < K; ++a) r[a] = div_exact(r[a], inv_cc);  // r / (1 + 2 lambda)
    float sum = r[0];
#pragma unroll
    for (int a = 1; a < K; ++a) sum += r[a];
#pragma unroll
    for (int a = 0; a < K; ++a) {
      const float f = r[a] / sum;
      changed |= __float_as_uint(f) != __float_as_uint(v[c][6][a]);
      st_wt(&gr.fout[a][gi], f);
    }
  }
}
// A chunk without FLUID in a full sweep. Interior cells are AIR or FLUID, so all of its
// cells are AIR: the pressure sweep writes 0 there (jacobi_value of an AIR cell) and a full
// diffusion sweep copies each cell's own fractions (diffusion_cell's AIR branch). The
// same stores, without the neighbour, content and right-hand-side loads.
template <int K>
__device__ __forceinline__ void air_chunk(const Geom& G, const PairedFields& gr, bool doJ, bool doD, int ch) {
  SweepCell q[kSweepCells];
  float v[kSweepCells][K];
#pragma unroll
  for (int c = 0; c < kSweepCells; ++c) {
    q[c] = sweep_cell(G, ch, c);
#pragma unroll
    for (int a = 0; a < K; ++a) v[c][a] = doD ? gr.fin[a][q[c].gi] : 0.f;
  }
#pragma unroll
  for (int c = 0; c < kSweepCells; ++c) {
    if (!q[c].ok) continue;
    if (doJ) st_wt(&gr.pout[q[c].gi], 0.f);
    if (doD) {
#pragma unroll
      for (int a = 0; a < K; ++a) st_wt(&gr.fout[a][q[c].gi], v[c][a]);
    }
  }
}
// Pressure sweep i paired with diffusion sweep i (`sweep` = i). A diffusion sweep that
// changes no bit leaves both fraction buffers equal (the skipped chunks are AIR,
// identical in both after the first sweep copied them), and every later sweep then
// recomputes the same values from the same inputs; so each sweep records whether it
// changed anything (gr.dchange[i % 3]) and a sweep whose predecessor changed nothing
// skips its diffusion part: the results are those of running it. Sweep i also clears
// the word sweep i + 1 will record into (word 0 is cleared by k_active_list, before
// each step's first sweep). Word 3 counts the paired sweeps that computed diffusion.
// Fields of a paired sweep past the ones the sweep's first loads need.
struct PairedAux {
  const uint8_t* content;
  const uint32_t* active;
  const float* fin[4];
  float* fout[4];
};
template <int K, bool PZERO>
__global__ void k_jacobi_diffusion(const uint4* __restrict__ tab, const uint32_t* __restrict__ list,
                                   uint32_t* __restrict__ dchange, const float* __restrict__ rhs,
                                   const float* __restrict__ pin, float* __restrict__ pout, int sweep,
                                   int mode, int nb, PairedAux aux, float lambda, double inv_cc, Geom G) {
  const bool fullJ = mode & 1, fullD = mode & 2, skip_ok = mode & 4;
  // Either sweep full: every chunk, each part where its own rule asks; otherwise the
  // active list, both parts. The previous sweep's change flag ((sweep + 2) % 3 is
  // (sweep - 1) % 3 from sweep 1 on), the table entry and the geometry are loaded together.
  // A listed sweep whose diffusion part is skipped then runs the pressure chunks in a loop
  // of its own, placed first, and never loads the diffusion part's arguments: 3.75 ->
  // 3.59 us per such sweep at 128^3 against 3.27 for a pressure sweep alone.
  const bool full = fullJ || fullD;
  const uint32_t prev = dchange[(sweep + 2) % 3];
  const uint4 e = tab[full ? 0u : blockIdx.x];
  const int nj = nb >= kXCD ? nb / kXCD : nb;
  AQF_SGPR_READY(G, e, "s"(prev), "s"(nj), "s"(nb));
  const bool skipD = skip_ok && !fullD && sweep > 0 && prev == 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st_wt(&dchange[(sweep + 1) % 3], 0u);
    if (!skipD) atomicAdd(&dchange[3], 1u);  // paired diffusion sweeps computed
  }
  if (__builtin_expect(skipD && !full, 1)) {  // the listed pressure sweep alone
    const int n = (int)e.x;
    for (int k = 0; k < n; ++k)
      jacobi_chunk<PZERO>(G, rhs, pin, pout, (int)(k == 0 ? e.z : k == 1 ? e.w : list[1 + e.y + k * nj]));
    return;
  }
  // The diffusion part's arguments, loaded together (by the sweeps that get here only).
  if (K == 1)
    asm volatile("" ::"s"(aux.content), "s"(aux.active), "s"(lambda), "s"(inv_cc), "s"(aux.fin[0]), "s"(aux.fout[0]));
  else if (K == 2)
    asm volatile("" ::"s"(aux.content), "s"(aux.active), "s"(lambda), "s"(inv_cc), "s"(aux.fin[0]), "s"(aux.fout[0]),
                 "s"(aux.fin[1]), "s"(aux.fout[1]));
  else if (K == 3)
    asm volatile("" ::"s"(aux.content), "s"(aux.active), "s"(lambda), "s"(inv_cc), "s"(aux.fin[0]), "s"(aux.fout[0]),
                 "s"(aux.fin[1]), "s"(aux.fout[1]), "s"(aux.fin[2]), "s"(aux.fout[2]));
  else
    asm volatile("" ::"s"(aux.content), "s"(aux.active), "s"(lambda), "s"(inv_cc), "s"(aux.fin[0]), "s"(aux.fout[0]),
                 "s"(aux.fin[1]), "s"(aux.fout[1]), "s"(aux.fin[2]), "s"(aux.fout[2]), "s"(aux.fin[3]),
                 "s"(aux.fout[3]));
  PairedFields f;
  f.rhs = rhs;
  f.pin = pin;
  f.pout = pout;
  f.content = aux.content;
  f.active = aux.active;
  for (int k = 0; k < 4; ++k) {
    f.fin[k] = aux.fin[k];
    f.fout[k] = aux.fout[k];
  }
  bool changed = false;
  auto body = [&](int ch) {
    const bool act = !full || chunk_active(f.active, ch);
    const bool doJ = fullJ || act, doD = fullD || act;
    if (!act) {  // a full sweep's chunk without FLUID (skipD implies !fullD: no copy then)
      air_chunk<K>(G, f, doJ, doD && !skipD, ch);
    } else if (skipD) {  // the pressure sweep alone: no diffusion loads
      if (doJ) jacobi_chunk<PZERO>(G, f.rhs, f.pin, f.pout, ch);
    } else if (doJ || doD) {
      paired_chunk<K, PZERO>(G, f, 0, 0, lambda, inv_cc, fullD, doJ, doD, ch, changed);
    }
  };
  if (full) {
    const int per = nb / kXCD, rem = nb % kXCD, bid = blockIdx.x;
    const int xcd = bid % kXCD, k = bid / kXCD;
    body(xcd * per + min(xcd, rem) + k);
  } else {
    const int n = (int)e.x;
    for (int k = 0; k < n; ++k) body((int)(k == 0 ? e.z : k == 1 ? e.w : list[1 + e.y + k * nj]));
  }
  // One plain store of 1 per wave that changed something: every writer stores the same
  // value, so the race is benign (a same-address atomic from every wave serialises: 39 us
  // per sweep in a mixing scene; so do same-address stores written through the L2, 38 us).
  if (__ballot(changed) && (threadIdx.x & 63) == 0) dchange[sweep % 3] = 1u;
}
template <int K>
void launch_paired(const Geom& G, Grid gr, int pin, int fullJ, int fin, float lambda, double inv_cc,
                   int fullD, int pzero, int sweep, int skip_ok, hipStream_t s) {
  const dim3 grid(fullJ || fullD ? sweep_grid(G, true) : gr.listed[1]);
  const uint4* tab = gr.btab + gr.listed[0];  // the paired grid's entries
  PairedAux aux;
  aux.content = gr.content;
  aux.active = gr.active;
  for (int a = 0; a < 4; ++a) {
    aux.fin[a] = gr.nvf[fin][a];
    aux.fout[a] = gr.nvf[fin ^ 1][a];
  }
  const int mode = (fullJ ? 1 : 0) | (fullD ? 2 : 0) | (skip_ok ? 4 : 0);
  const int nb = (int)grid.x;
  if (pzero)
    hipLaunchKernelGGL((k_jacobi_diffusion<K, true>), grid, dim3(kBlock), 0, s, tab, gr.alist, gr.dchange, gr.rhs,
                       gr.p[pin], gr.p[pin ^ 1], sweep, mode, nb, aux, lambda, inv_cc, G);
  else
    hipLaunchKernelGGL((k_jacobi_diffusion<K, false>), grid, dim3(kBlock), 0, s, tab, gr.alist, gr.dchange, gr.rhs,
                       gr.p[pin], gr.p[pin ^ 1], sweep, mode, nb, aux, lambda, inv_cc, G);
}
void launch_jacobi_diffusion(const Geom& G, Grid gr, int K, int pin, int fullJ, int fin, float lambda,
                             double inv_cc, int fullD, int pzero, int sweep, int skip_ok, hipStream_t s) {
  switch (K) {
    case 1: launch_paired<1>(G, gr, pin, fullJ, fin, lambda, inv_cc, fullD, pzero, sweep, skip_ok, s); break;
    case 2: launch_paired<2>(G, gr, pin, fullJ, fin, lambda, inv_cc, fullD, pzero, sweep, skip_ok, s); break;
    case 3: launch_paired<3>(G, gr, pin, fullJ, fin, lambda, inv_cc, fullD, pzero, sweep, skip_ok, s); break;
    default: launch_paired<4>(G, gr, pin, fullJ, fin, lambda, inv_cc, fullD, pzero, sweep, skip_ok, s); break;
  }
}

// ------------------------------------------------------------------ G2P and advection
// getInterpolatedValueAtPoint (Fluid/MAC_Grid_3D.cuh:66-97) on one face component:
// clamp to [0, N-1], trilinear lerps in the order z, y, x.
__device__ __forceinline__ float interp(const Geom& G, const float* __restrict__ F, float x, float y,
                                        float z) {
  x = fmaxf(fminf(x, (float)G.X - 1.f), 0.f);
  y = fmaxf(fminf(y, (float)G.Y - 1.f), 0.f);
  z = fmaxf(fminf(z, (float)G.Z - 1.f), 0.f);
  const int i = (int)floorf(x), j = (int)floorf(y), k = (int)floorf(z);
  const float tx = x - (float)i, ty = y - (float)j, tz = z - (float)k;
  const long long b = G.g(i, j, k);
  const long long sx = G.sx, sy = G.EZ;
  const float x0y0 = lerpf_(F[b], F[b + 1], tz);
  const float x0y1 = lerpf_(F[b + sy], F[b + sy + 1], tz);
  const float x1y0 = lerpf_(F[b + sx], F[b + sx + 1], tz);
  const float x1y1 = lerpf_(F[b + sx + sy], F[b + sx + sy + 1], tz);
  const float x0 = lerpf_(x0y0, x0y1, ty);
  const float x1 = lerpf_(x1y0, x1y1, ty);
  return lerpf_(x0, x1, tx);
}
// getPointVelocity / getPointNewVelocity (Fluid/MAC_Grid_3D.cuh:100-130).
__device__ __forceinline__ void point_velocity(const Geom& G, const float* U, const float* V,
                                               const float* W, float px, float py, float pz,
                                               float out[3]) {
  const float x = div_exact(px, G.inv_h), y = div_exact(py, G.inv_h), z = div_exact(pz, G.inv_h);
  out[0] = interp(G, U, x, y - 0.5f, z - 0.5f);
  out[1] = interp(G, V, x - 0.5f, y, z - 0.5f);
  out[2] = interp(G, W, x - 0.5f, y - 0.5f, z);
}

// One particle record in registers.
struct PRec {
  float p[3], v[3], f[4];
};
template <int K>
__device__ __forceinline__ void load_rec(const Particles& P, int i, PRec& r) {
  const float4 a = P.a[i], b = P.b[i];
  r.p[0] = a.x, r.p[1] = a.y, r.p[2] = a.z, r.v[0] = a.w;
  r.v[1] = b.x, r.v[2] = b.y, r.f[0] = b.z, r.f[1] = b.w;
  r.f[2] = r.f[3] = 0.f;
  if (K > 2) {
    const float4 c = P.c[i];
    r.f[2] = c.x, r.f[3] = c.y;
  }
}
template <int K>
__device__ __forceinline__ void store_rec(const Particles& P, int i, const PRec& r) {
  P.a[i] = make_float4(r.p[0], r.p[1], r.p[2], r.v[0]);
  P.b[i] = make_float4(r.v[1], r.v[2], r.f[0], r.f[1]);
  if (K > 2) P.c[i] = make_float4(r.f[2], r.f[3], 0.f, 0.f);
}

// transferToParticlesImpl (Fluid/Fluid_3D_FLIP.cu:189-227). Returns newV at the
// particle, which is also RK3's first stage.
// interp on the interleaved (new, old) copy: the same corners, lerps and order for both
// fields, each corner pair of both in one 16-byte load.
__device__ __forceinline__ void interp2(const Geom& G, const float2* __restrict__ F, float x, float y, float z,
                                        float& nout, float& oout) {
  x = fmaxf(fminf(x, (float)G.X - 1.f), 0.f);
  y = fmaxf(fminf(y, (float)G.Y - 1.f), 0.f);
  z = fmaxf(fminf(z, (float)G.Z - 1.f), 0.f);
  const int i = (int)floorf(x), j = (int)floorf(y), k = (int)floorf(z);
  const float tx = x - (float)i, ty = y - (float)j, tz = z - (float)k;
  const long long b = G.g(i, j, k);
  const long long sx = G.sx, sy = G.EZ;
  struct alignas(8) Pair2 {  // faces k and k + 1 of a z row: (new, old) each
    float n0, o0, n1, o1;
  };
  auto pair = [&](long long o) {
    Pair2 v;
    __builtin_memcpy(&v, F + o, sizeof v);
    return v;
  };
  const Pair2 a = pair(b), c = pair(b + sy), d = pair(b + sx), e = pair(b + sx + sy);
  const float n00 = lerpf_(a.n0, a.n1, tz), n01 = lerpf_(c.n0, c.n1, tz);
  const float n10 = lerpf_(d.n0, d.n1, tz), n11 = lerpf_(e.n0, e.n1, tz);
  const float o00 = lerpf_(a.o0, a.o1, tz), o01 = lerpf_(c.o0, c.o1, tz);
  const float o10 = lerpf_(d.o0, d.o1, tz), o11 = lerpf_(e.o0, e.o1, tz);
  nout = lerpf_(lerpf_(n00, n01, ty), lerpf_(n10, n11, ty), tx);
  oout = lerpf_(lerpf_(o00, o01, ty), lerpf_(o10, o11, ty), tx);
}
__device__ __forceinline__ void point_velocity2(const Geom& G, const float2* const NO[3], float px, float py,
                                                float pz, float nvp[3], float ovp[3]) {
  const float x = div_exact(px, G.inv_h), y = div_exact(py, G.inv_h), z = div_exact(pz, G.inv_h);
  interp2(G, NO[0], x, y - 0.5f, z - 0.5f, nvp[0], ovp[0]);
  interp2(G, NO[1], x - 0.5f, y, z - 0.5f, nvp[1], ovp[1]);
  interp2(G, NO[2], x - 0.5f, y - 0.5f, z, nvp[2], ovp[2]);
}

// interleaved: new and old velocity from gr.nvo (valid at the start of a step);
// otherwise from newVelocity buffer nv and velocity.
template <int K>
__device__ __forceinline__ void g2p_one(const Geom& G, PRec& r, const Grid& gr, int nv, int nf,
                                        float flip, float nvOut[3], bool interleaved = false) {
  float nvp[3], ovp[3];
  if (interleaved) {
    const float2* const NO[3] = {gr.nvo[0], gr.nvo[1], gr.nvo[2]};
    point_velocity2(G, NO, r.p[0], r.p[1], r.p[2], nvp, ovp);
  } else {
    point_velocity(G, gr.nvel[nv][0], gr.nvel[nv][1], gr.nvel[nv][2], r.p[0], r.p[1], r.p[2], nvp);
    point_velocity(G, gr.vel[0], gr.vel[1], gr.vel[2], r.p[0], r.p[1], r.p[2], ovp);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float change = nvp[c] - ovp[c];
    const float fl = r.v[c] + change;
    const float pic = nvp[c];
    r.v[c] = flip * fl + (1.0f - flip) * pic;
    nvOut[c] = nvp[c];
  }
  const int cx = clampi((int)floorf(div_exact(r.p[0], G.inv_h)), 0, G.X - 1);
  const int cy = clampi((int)floorf(div_exact(r.p[1], G.inv_h)), 0, G.Y - 1);
  const int cz = clampi((int)floorf(div_exact(r.p[2], G.inv_h)), 0, G.Z - 1);
  const long long gc = G.g(cx, cy, cz);
  float f[K];
#pragma unroll
  for (int a = 0; a < K; ++a) f[a] = gr.nvf[nf][a][gc];
  float s = f[0] * 1.f;
#pragma unroll
  for (int a = 1; a < K; ++a) s += f[a] * 1.f;
  if (s > 0.f) {
#pragma unroll
    for (int a = 0; a < K; ++a) f[a] /= s;
  }
#pragma unroll
  for (int a = 0; a < K; ++a) r.f[a] = f[a];
}

// moveParticlesImpl (Fluid/Fluid_3D_FLIP.cu:229-285): Ralston RK3 on newVelocity,
// clamp to [dx/4, L - dx/4] with velocity reflection on clamp.
__device__ __forceinline__ void advect_one(const Geom& G, PRec& r, const Grid& gr, int nv, float dt,
                                           const float u1[3]) {
  const float b[3] = {r.p[0], r.p[1], r.p[2]};
  const float* U = gr.nvel[nv][0];
  const float* V = gr.nvel[nv][1];
  const float* W = gr.nvel[nv][2];
  float u2[3], u3[3], q[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) q[c] = b[c] + (dt * u1[c]) / 2.f;
  point_velocity(G, U, V, W, q[0], q[1], q[2], u2);
#pragma unroll
  for (int c = 0; c < 3; ++c) q[c] = b[c] + ((dt * u2[c]) * 3.f) / 4.f;
  point_velocity(G, U, V, W, q[0], q[1], q[2], u3);
  const float m = G.h / 4.f;
  const float L[3] = {(float)G.X * G.h, (float)G.Y * G.h, (float)G.Z * G.h};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double ninth = 1.0 / 9.0;
    float d = b[c] + dt * ((div_exact(u1[c] * 2.f, ninth) + div_exact(u2[c] * 3.f, ninth)) +
                           div_exact(u3[c] * 4.f, ninth));
    if (d < m) {
      d = m;
      r.v[c] *= -1.f;
    }
    if (d > L[c] - m) {
      d = L[c] - m;
      r.v[c] *= -1.f;
    }
    r.p[c] = d;
  }
}

template <int K>
__global__ void k_g2p(Geom G, Particles P, long long Np, Grid gr, int nv, int nf, float flip) {
  const int i = gtid();
  if (i >= Np) return;
  PRec r;
  load_rec<K>(P, i, r);
  float u1[3];
  g2p_one<K>(G, r, gr, nv, nf, flip, u1);
  store_rec<K>(P, i, r);
}

__global__ void k_advect(Geom G, Particles P, long long Np, Grid gr, int nv, float dt) {
  const int i = gtid();
  if (i >= Np) return;
  PRec r;
  load_rec<2>(P, i, r);
  float u1[3];
  point_velocity(G, gr.nvel[nv][0], gr.nvel[nv][1], gr.nvel[nv][2], r.p[0], r.p[1], r.p[2], u1);
  advect_one(G, r, gr, nv, dt, u1);
  P.a[i] = make_float4(r.p[0], r.p[1], r.p[2], r.v[0]);
  P.b[i] = make_float4(r.v[1], r.v[2], r.f[0], r.f[1]);
}

// Fused G2P + RK3 + hash + binning: one pass over the particles per step.
// Two particles per thread, kBlock apart (loads stay coalesced): both records are loaded
// first, so the second one's load is in flight while the first particle runs its
// dependent gather rounds, and the launch has half the waves. The empty asm statements
// keep the compiler from hoisting the second particle's gathers into the first one's
// rounds (that took 128 VGPRs and spilled). G2P + RK3 + hash 146 -> 142 us at 128^3.
// The two particles are binned by separate calls, so a cell can receive more runs of
// ranks than with one particle per thread; k_segsort orders every such cell.
template <int K>
__global__ void k_g2p_advect_hash(Geom G, Particles P, long long Np, Grid gr, int nv, int nf,
                                  float flip, float dt, int* __restrict__ hash,
                                  int* __restrict__ rank, int* __restrict__ count,
                                  uint8_t* __restrict__ multi) {
  const int i0 = xcd_block() * (2 * kBlock) + (int)threadIdx.x, i1 = i0 + kBlock;
  if (i0 == 0) *gr.stepstamp += 1u;  // the step's stamp: no other thread of this pass reads it
  const bool v0 = i0 < Np, v1 = i1 < Np;
  PRec r0, r1;
  load_rec<K>(P, v0 ? i0 : 0, r0);
  load_rec<K>(P, v1 ? i1 : 0, r1);  // (index 0 for a lane past the end: loaded, never stored)
  asm volatile("" ::: "memory");
  float a0[3];
  g2p_one<K>(G, r0, gr, nv, nf, flip, a0, true);
  advect_one(G, r0, gr, nv, dt, a0);
  if (v0) store_rec<K>(P, i0, r0);
  const int h0 = v0 ? cell_hash(G, r0.p[0], r0.p[1], r0.p[2]) : 0;
  asm volatile("" ::: "memory");
  float a1[3];
  g2p_one<K>(G, r1, gr, nv, nf, flip, a1, true);
  advect_one(G, r1, gr, nv, dt, a1);
  if (v1) store_rec<K>(P, i1, r1);
  const int h1 = v1 ? cell_hash(G, r1.p[0], r1.p[1], r1.p[2]) : 0;
  int k0, k1;
  bin_particle(h0, v0, count, multi, k0);
  bin_particle(h1, v1, count, multi, k1);
  if (v0) {
    hash[i0] = h0;
    rank[i0] = k0;
  }
  if (v1) {
    hash[i1] = h1;
    rank[i1] = k1;
  }
}

void launch_g2p(const Geom& G, Particles P, long long Np, int K, Grid gr, int nv, int nf,
                StepParams sp, hipStream_t s) {
  const int nb = nblocks(Np);
  switch (K) {
    case 1: hipLaunchKernelGGL(k_g2p<1>, dim3(nb), dim3(kBlock), 0, s, G, P, Np, gr, nv, nf, sp.flip); break;
    case 2: hipLaunchKernelGGL(k_g2p<2>, dim3(nb), dim3(kBlock), 0, s, G, P, Np, gr, nv, nf, sp.flip); break;
    case 3: hipLaunchKernelGGL(k_g2p<3>, dim3(nb), dim3(kBlock), 0, s, G, P, Np, gr, nv, nf, sp.flip); break;
    default: hipLaunchKernelGGL(k_g2p<4>, dim3(nb), dim3(kBlock), 0, s, G, P, Np, gr, nv, nf, sp.flip); break;
  }
}
void launch_advect(const Geom& G, Particles P, long long Np, Grid gr, int nv, StepParams sp,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_advect, dim3(nblocks(Np)), dim3(kBlock), 0, s, G, P, Np, gr, nv, sp.dt);
}
void launch_g2p_advect_hash(const Geom& G, Particles P, long long Np, int K, Grid gr, int nv,
                            int nf, StepParams sp, Bins B, hipStream_t s) {
  const int nb = (int)((Np + 2 * kBlock - 1) / (2 * kBlock));
  switch (K) {
    case 1: hipLaunchKernelGGL(k_g2p_advect_hash<1>, dim3(nb), dim3(kBlock), 0, s, G, P, Np, gr, nv, nf, sp.flip, sp.dt, B.hash, B.rank, B.count, B.multi); break;
    case 2: hipLaunchKernelGGL(k_g2p_advect_hash<2>, dim3(nb), dim3(kBlock), 0, s, G, P, Np, gr, nv, nf, sp.flip, sp.dt, B.hash, B.rank, B.count, B.multi); break;
    case 3: hipLaunchKernelGGL(k_g2p_advect_hash<3>, dim3(nb), dim3(kBlock), 0, s, G, P, Np, gr, nv, nf, sp.flip, sp.dt, B.hash, B.rank, B.count, B.multi); break;
    default: hipLaunchKernelGGL(k_g2p_advect_hash<4>, dim3(nb), dim3(kBlock), 0, s, G, P, Np, gr, nv, nf, sp.flip, sp.dt, B.hash, B.rank, B.count, B.multi); break;
  }
}

// ------------------------------------------------------------------ FP64 reductions
// Deterministic two-level reductions: fixed-shape block trees write per-block
// partials, one block folds the partials in index order.
__device__ double block_sum_d(double v) {
  __shared__ double sh[kBlock];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// Residual of the pressure system over FLUID cells: A has diagonal
// 6 - #(out-of-domain neighbours) and -1 per FLUID neighbour, AIR neighbours are
// 0 (Fluid/Legacy/PressureEquation.cuh:216-328); b = -div*density.
// Per chunk (kSweepSpan cells, two per thread, every load issued before use) the
// sums of r^2 and b^2 over its FLUID cells, folded by a fixed-shape block tree into
// partial[2*chunk + 0/1]; chunks without FLUID write zeros. The fold over the chunks
// (k_fold_partials) runs in a fixed order, so the result is deterministic.
__global__ void k_residual_partial(Geom G, const uint8_t* __restrict__ content,
                                   const float* __restrict__ rhs, const float* __restrict__ p,
                                   const uint32_t* __restrict__ active, double* __restrict__ partial) {
  const int ch = xcd_block();
  double rr = 0.0, bb = 0.0;
  if (chunk_active(active, ch)) {  // uniform per workgroup
    SweepCell q[kSweepCells];
    uint8_t c0[kSweepCells], cn[kSweepCells][6];
    float bv[kSweepCells], pc[kSweepCells], pn[kSweepCells][6];
#pragma unroll
    for (int c = 0; c < kSweepCells; ++c) {
      q[c] = sweep_cell(G, ch, c);
      c0[c] = content[q[c].gi];
      bv[c] = rhs[q[c].gi];
      pc[c] = p[q[c].gi];
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        cn[c][k] = content[q[c].nb[k]];
        pn[c][k] = p[q[c].nb[k]];
      }
    }
#pragma unroll
    for (int c = 0; c < kSweepCells; ++c) {
      if (!q[c].ok || c0[c] != FLUID) continue;
      // A: diagonal 6 minus one per out-of-domain neighbour (a clamped neighbour index
      // equals the cell's own), -1 per FLUID neighbour, AIR neighbours 0.
      const double x = (double)pc[c];
      double ap = 6.0 * x;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        if (q[c].nb[k] == q[c].gi)
          ap -= x;
        else if (cn[c][k] == FLUID)
          ap -= (double)pn[c][k];
      }
      const double b = (double)bv[c];
      const double r = b - ap;
      rr += r * r;
      bb += b * b;
    }
  }
  rr = block_sum_d(rr);
  bb = block_sum_d(bb);
  if (threadIdx.x == 0) {
    partial[2 * ch] = rr;
    partial[2 * ch + 1] = bb;
  }
}

__global__ void k_fold_partials(const double* __restrict__ partial, int nparts, int width,
                                double* __restrict__ out) {
  // One block; thread t folds column t % width over partial rows t/width, t/width+..
  // then a fixed-order tree over the rows of each column.
  __shared__ double sh[kBlock];
  const int rows = kBlock / width;
  const int col = threadIdx.x % width, row = threadIdx.x / width;
  double s = 0.0;
  if (row < rows)
    for (int r = row; r < nparts; r += rows) s += partial[(long long)r * width + col];
  sh[threadIdx.x] = s;
  __syncthreads();
  if ((int)threadIdx.x < width) {
    double t = 0.0;
    for (int r = 0; r < rows; ++r) t += sh[r * width + threadIdx.x];
    out[threadIdx.x] = t;
  }
}

// Fold of the residual's per-chunk pairs: 1024 threads, thread t adds pairs t,
// t + 1024, ... in order, then a fixed-shape tree (deterministic).
__global__ __launch_bounds__(1024) void k_fold_pairs_1024(const double* __restrict__ partial, int n,
                                                          double* __restrict__ out) {
  __shared__ double s1[1024], s2[1024];
  double a = 0.0, b = 0.0;
  for (int i = (int)threadIdx.x; i < n; i += 1024) {
    a += partial[2 * i];
    b += partial[2 * i + 1];
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s1[0];
    out[1] = s2[0];
  }
}
void launch_residual(const Geom& G, Grid gr, int pb, double* partial, double* out, hipStream_t s) {
  const int nch = sweep_chunks(G);
  hipLaunchKernelGGL(k_residual_partial, dim3(nch), dim3(kBlock), 0, s, G, gr.content, gr.rhs,
                     gr.p[pb], gr.active, partial);
  hipLaunchKernelGGL(k_fold_pairs_1024, dim3(1), dim3(1024), 0, s, partial, nch, out);
}

// Volume and mass checks: per phase the sum of particle fractions, the number of
// FLUID cells and the particle position bounds (min stored negated so every
// column folds with a sum or a max).
constexpr int kStatsWidth = 16;  // f0..f3, fluid cells, -min xyz, max xyz, pad
__global__ void k_stats_partial(Geom G, Particles P, long long Np, int K,
                                const uint8_t* __restrict__ content, double* __restrict__ partial) {
  double f[4] = {0, 0, 0, 0};
  double fluid = 0;
  float mn[3] = {3.4e38f, 3.4e38f, 3.4e38f}, mx[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < Np; i += stride) {
    const float4 a = P.a[i], b = P.b[i];
    f[0] += (double)b.z;
    f[1] += (double)b.w;
    if (K > 2) {
      const float4 c = P.c[i];
      f[2] += (double)c.x;
      f[3] += (double)c.y;
    }
    const float q[3] = {a.x, a.y, a.z};
#pragma unroll
    for (int a3 = 0; a3 < 3; ++a3) {
      mn[a3] = fminf(mn[a3], q[a3]);
      mx[a3] = fmaxf(mx[a3], q[a3]);
    }
  }
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < G.ncells_hash; n += stride) {
    int x, y, z;
    cell_of(G, n, x, y, z);
    if (content[G.g(x, y, z)] == FLUID) fluid += 1.0;
  }
  double col[kStatsWidth];
#pragma unroll
  for (int a = 0; a < 4; ++a) col[a] = block_sum_d(f[a]);
  col[4] = block_sum_d(fluid);
  __shared__ float shm[kBlock];
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    shm[threadIdx.x] = a < 3 ? -mn[a] : mx[a - 3];
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) shm[threadIdx.x] = fmaxf(shm[threadIdx.x], shm[threadIdx.x + o]);
      __syncthreads();
    }
    col[5 + a] = (double)shm[0];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int a = 0; a < 11; ++a) partial[(long long)blockIdx.x * kStatsWidth + a] = col[a];
}

__global__ void k_fold_stats(const double* __restrict__ partial, int nparts, double* __restrict__ out) {
  if (threadIdx.x >= kStatsWidth) return;
  const int c = threadIdx.x;
  if (c >= 11) {
    out[c] = 0;
    return;
  }
  double s = (c < 5) ? 0.0 : -1.0e300;
  for (int r = 0; r < nparts; ++r) {
    const double v = partial[(long long)r * kStatsWidth + c];
    s = (c < 5) ? s + v : fmax(s, v);
  }
  out[c] = s;
}

void launch_stats(const Geom& G, Particles P, long long Np, int K, Grid gr, double* partial,
                  int nparts, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_stats_partial, dim3(nparts), dim3(kBlock), 0, s, G, P, Np, K, gr.content,
                     partial);
  hipLaunchKernelGGL(k_fold_stats, dim3(1), dim3(64), 0, s, partial, nparts, out);
}

}  // namespace aqf
