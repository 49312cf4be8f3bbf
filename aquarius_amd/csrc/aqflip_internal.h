// aqflip_internal.h -- device data layout and kernel launchers of libaqflip.
//
// HBM layout (one GPU, everything resident for the life of the handle):
//   * Particles: 16-byte records (px,py,pz,vx)(vy,vz,f0,f1)[(f2,f3,-,-)], i.e. 32
//     bytes per particle with up to 2 phases (48 with 3-4), two buffers (current
//     and sorted copy). Every particle kernel moves whole records with 16-byte
//     loads; only the first K fraction channels are used (K = phases in use), the
//     rest stay exactly zero.
//   * Grid: every field uses one padded cell index over the face extent
//     (X+1)x(Y+1)x(Z+1) of MAC_Grid_3D (Fluid/MAC_Grid_3D.cu:53-71):
//     g(x,y,z) = (x*EY + y)*EZ + z with EY = Y+1 and EZ = Z+1 rounded up to 8, so
//     z rows are 32-byte aligned and all fields share stencil offsets (sx, sy, 1).
//     Cell-centred fields (pressure, divergence, density, fractions) use the
//     interior x<X, y<Y, z<Z only. Face velocities are split per component
//     (u, v, w) so each stencil reads one dense array.
//   * Cell bins: counts, exclusive offsets (X*Y*Z + 1) over the particle hash
//     x*Y*Z + y*Z + z (Fluid/Fluid_3D_kernels.cuh:8-27), per-particle hash and
//     in-cell rank, and the sort permutation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aqf {

enum : uint8_t { AIR = 0, FLUID = 1, SOLID = 2 };  // Fluid/MAC_Grid_3D.cuh:15-17

// Division by a runtime constant d (n < 2^31): q = (umulhi(n, m) + n) >> s with
// s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1.
struct FastDiv {
  uint32_t d, m, s;
};
inline FastDiv make_fastdiv(uint32_t d) {
  uint32_t s = 0;
  while ((1ull << s) < d) ++s;
  const uint64_t m = ((1ull << 32) * ((1ull << s) - d)) / d + 1;
  return FastDiv{d, (uint32_t)m, s};
}
__host__ __device__ inline uint32_t fast_div(uint32_t n, FastDiv f) {
#ifdef __HIP_DEVICE_COMPILE__
  const uint32_t t = __umulhi(n, f.m);
#else
  const uint32_t t = (uint32_t)(((uint64_t)n * f.m) >> 32);
#endif
  return (t + n) >> f.s;
}

struct Geom {
  int X, Y, Z;        // cells
  int EY, EZ;         // padded strides: EY = Y+1, EZ = round_up(Z+1, 8)
  long long sx;       // x stride = (long long)EY*EZ
  long long ncell;    // (X+1)*EY*EZ
  int ncells_hash;    // X*Y*Z
  int nfaces;         // (X+1)*(Y+1)*(Z+1)
  float h;            // cellPhysicalSize
  double inv_h;       // 1/h rounded to double (see div_exact)
  FastDiv divZ, divYZ, divZ1, divYZ1;  // Z, Y*Z, Z+1, (Y+1)*(Z+1)
  __host__ __device__ long long g(int x, int y, int z) const {
    return (long long)x * sx + (long long)y * EZ + z;
  }
};

// Particle state as 16-byte records, in the current (sorted) order:
//   a = (px, py, pz, vx), b = (vy, vz, f0, f1), c = (f2, f3, 0, 0) (c only with K > 2).
struct Particles {
  float4* a;
  float4* b;
  float4* c;
};

// Bit pattern of the Jacobi right-hand side of an AIR cell: a NaN payload that
// arithmetic never produces (AMD and x86 return the canonical 0x7FC00000).
constexpr uint32_t kAirRhsBits = 0x7FA1A1A1u;

struct Grid {
  uint8_t* content;     // AIR/FLUID/SOLID
  float* vel[3];        // velocity (P2G result, pre-gravity)
  float* nvel[2][3];    // newVelocity, ping-pong for extrapolation
  uint8_t* hasvel[2];   // hasVelocity bits 0..2, ping-pong
  float* p[2];          // pressure, ping-pong
  float* div;           // divergence
  float* rhs;           // -divergence*density (Jacobi right-hand side); kAirRhsBits on AIR
  float* density;
  float* vf[4];         // volumeFractions
  float* nvf[2][4];     // newVolumeFractions, ping-pong
  uint32_t* active;     // per 256-cell launch block of the interior: any FLUID cell
                        // (32-bit so a block reads its flag with one scalar load)
  uint32_t* alist;      // [0] = count, then the ids of the sweep chunks holding FLUID
  uint4* btab;          // per workgroup of a listed sweep grid: {chunks, first list
                        // position, first two chunk ids}; the pressure grid's entries,
                        // then the paired grid's (launch_active_list writes both)
  int listed[2];        // workgroups of a listed pressure sweep and of a listed paired sweep
  uint32_t* dchange;    // words 0-2: paired diffusion sweep i changed a bit -> word i % 3 != 0;
                        // word 3: paired diffusion sweeps computed (not skipped)
  uint32_t* stepstamp;  // one word, advanced by the G2P pass that starts every step
  uint32_t* rowstamp;   // per face row (x, y), x <= X, y <= Y: the stamp of the last step
                        // whose P2G found FLUID in a cell row within 4 rows in x and in y
  float2* nvo[3];       // per face: (newVelocity, velocity) interleaved, for the G2P sample
                        // of the next step (one 16-byte load per corner pair, both fields);
                        // written by a step's last extrapolation pass on the rows it visits,
                        // or in full by launch_build_nvo
};

struct Bins {
  int* count;     // X*Y*Z per-cell counts
  int* offsets;   // X*Y*Z + 1 exclusive prefix
  int* partials;  // scan block partials
  int* hash;      // per particle, previous order
  int* rank;      // per particle: slot within its cell's segment (pre-sort)
  int* perm;      // sorted slot -> previous index
  uint8_t* multi; // X*Y*Z: 1 when a cell received more than one run of ranks (its segment
                  // may be out of order); k_segsort sorts those and clears every flag
};

struct StepParams {
  float dt, flip;
  float g[3];
  float rho[4];
  float lambda;  // D*dt/dx^2
  double inv_cc; // 1/(1 + 2*lambda) rounded to double
  int coverX, coverY, coverZ;  // fixBoundary pair coverage
};

// One seeding volume as the GPU seeder walks it (seed_kernels.hip): cells in the
// host's loop order over extents n[0..2] from origin o[0..2], loop level j running
// along grid axis axis[j]; draws start at stream index `base`.
struct SeedVolume {
  int n[3], o[3], axis[3];
  int sphere;
  float cc[3], limit;
  int phase;
  uint64_t base;
};
struct SeedParams {
  uint64_t seed;
  float h, jitter;
  int S[3];
};

// ---- launchers (flip_kernels.hip, seed_kernels.hip) ----
void launch_fill_u32(void* p, uint32_t v, long long n, hipStream_t s);
void launch_seed_count(const SeedVolume& v, const SeedParams& sp, long long nc, int* counts, hipStream_t s);
void launch_seed_write(const SeedVolume& v, const SeedParams& sp, long long nc, const int* offsets,
                       long long out0, Particles P, hipStream_t s);
int seed_block_size();
void launch_hash_count(const Geom& G, Particles P, long long Np, Bins B, hipStream_t s);
void launch_scan(const Geom& G, Bins B, hipStream_t s);
void launch_scatter_sort_gather(const Geom& G, Particles cur, Particles alt, long long Np, int K,
                                Bins B, hipStream_t s);
// Standalone spatial hashing (records of `stride` floats, position first).
void launch_hash_records(const Geom& G, const float* pos, long long n, int stride, Bins B, hipStream_t s);
void launch_scatter_segsort(const Geom& G, long long n, Bins B, hipStream_t s);
void launch_hash_outputs(const Geom& G, long long n, int stride, Bins B, const float* rec, int* start, int* end,
                         int* sorted_hash, float* sorted_rec, hipStream_t s);
void launch_p2g(const Geom& G, Particles P, long long Np, int K, Bins B, Grid gr, int nv,
                int hv, int nf, StepParams sp, hipStream_t s);
void launch_gravity_boundary(const Geom& G, Grid gr, int nv, StepParams sp, int doGravity,
                             int doBoundary, hipStream_t s);
void launch_divergence(const Geom& G, Grid gr, int nv, Bins B, hipStream_t s);
void launch_refresh_rhs(const Geom& G, Grid gr, hipStream_t s);
// Rebuild gr.alist from gr.active (both launchers above end with it).
void launch_active_list(const Geom& G, Grid gr, hipStream_t s);
// Number of sweep chunks (the unit of gr.alist).
int sweep_chunks(const Geom& G);
// Entries of gr.btab (one per workgroup of both listed sweep grids, at their largest).
int sweep_table_entries(const Geom& G);
// gr.listed from the default grid sizes, or at most `limit` workgroups (limit > 0: a
// diagnostic that makes each workgroup walk many list entries).
void set_listed_grids(const Geom& G, Grid& gr, int limit);
void launch_fill_pressure_zero(const Geom& G, Grid gr, int pb, hipStream_t s);
// full = 1: every cell written (AIR cells get 0); full = 0: blocks without FLUID
// cells and AIR cells are skipped -- valid once both buffers hold 0 on AIR.
// pzero = 1: the input is the zero guess (not read); needs full = 1.
void launch_jacobi(const Geom& G, Grid gr, int pin, int full, int pzero, hipStream_t s);
void launch_air_pressure_zero(const Geom& G, Grid gr, int pb, hipStream_t s);
void launch_project(const Geom& G, Grid gr, int nv, int hv, int pb, hipStream_t s);
// near_rows: only the face rows stamped by this step's P2G (passes 2 and 3 of a step, see
// k_extrapolate); otherwise every face.
void launch_extrapolate(const Geom& G, Grid gr, int nvin, int hvin, bool near_rows, hipStream_t s,
                        bool write_nvo = false);
// gr.nvo from newVelocity buffer nv and velocity, every face.
void launch_build_nvo(const Geom& G, Grid gr, int nv, hipStream_t s);
// full = 1: AIR cells copied to the output buffer; full = 0: AIR cells and blocks
// without FLUID skipped -- valid once both buffers agree on AIR cells.
void launch_diffusion(const Geom& G, Grid gr, int K, int fin, float lambda, double inv_cc, int full,
                      hipStream_t s);
// One Jacobi sweep (p[pin] -> p[pin^1]) and one diffusion sweep (nvf[fin] ->
// nvf[fin^1]) in one launch; fullJ / fullD as `full` above.
// sweep = index of the pair within the step (0 first); skip_ok = 1 lets a diffusion
// sweep whose predecessor changed nothing skip its (identical) work.
void launch_jacobi_diffusion(const Geom& G, Grid gr, int K, int pin, int fullJ, int fin, float lambda,
                             double inv_cc, int fullD, int pzero, int sweep, int skip_ok, hipStream_t s);
void launch_g2p(const Geom& G, Particles P, long long Np, int K, Grid gr, int nv, int nf,
                StepParams sp, hipStream_t s);
void launch_advect(const Geom& G, Particles P, long long Np, Grid gr, int nv, StepParams sp,
                   hipStream_t s);
void launch_g2p_advect_hash(const Geom& G, Particles P, long long Np, int K, Grid gr, int nv,
                            int nf, StepParams sp, Bins B, hipStream_t s);
// FP64 reductions: residual terms (sum r^2, sum b^2) and volume statistics.
void launch_residual(const Geom& G, Grid gr, int pb, double* partial, double* out, hipStream_t s);
void launch_stats(const Geom& G, Particles P, long long Np, int K, Grid gr, double* partial,
                  int nparts, double* out, hipStream_t s);

}  // namespace aqf
