// flip_oracle.hpp -- CPU oracle for the FLIP simulation step.
//
// TEST INFRASTRUCTURE ONLY. This is the checker the GPU library is compared
// against; it is never linked into, loaded by or called from the product
// (aquarius_amd / libaqflip). Only tests/, the smoke check and bench.py's CPU
// baseline leg use it.
//
// It restates the logic of the reference kernels of AmesingFlank/Aquarius
// (paths relative to that tree) as plain C++17, templated on the scalar type:
// Oracle<float> is the FP32 parity reference, Oracle<double> the FP64 run every
// result states its error against. Each function cites the file:line it follows.
//
// Deliberate, documented deviations from the reference (identical in the GPU
// library):
//   * Jacobi, velocity extrapolation and diffusion read the previous iterate and
//     write a new buffer (ping-pong). The reference updates in place while reading
//     neighbours through a texture in the same launch, which is formally undefined
//     (Fluid/Fluid_3D_kernels.cu:460-505, 210-422, 520-556); ping-pong is the
//     Jacobi iteration x_{i+1} = D^-1 (b - C x_i) of report/ChapterImplementFLIP.tex:297-302.
//   * IEEE single-precision division and no FMA contraction; the reference builds
//     with -use_fast_math (CMakeLists.txt:16).
//   * Ties of the sort keep the previous order (stable sort by hash); thrust's
//     sort_by_key (Fluid/Fluid_3D_common.cuh:35) does not document stability.
//   * Random jitter comes from a seeded SplitMix64 stream instead of rand().
//   * If gravity had more than one non-zero component, applyGravityImpl would race
//     on AIR cells written by several neighbours; contributions are summed in the
//     order x-, y-, z-neighbour (the order of a sequential sweep).
//   * Texture reads clamp to edge on every axis (addressMode[1..2] are left at the
//     wrap enum with unnormalised coordinates, Fluid/VolumeData.cuh:58-67).
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/aqflip.h"

namespace aqo {

enum { AIR = 0, FLUID = 1, SOLID = 2 };  // Fluid/MAC_Grid_3D.cuh:15-17

// Seeded replacement for random0to1() (Common/Commons.h:3-5): SplitMix64,
// 24 high bits -> float in [0,1).
struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  float u01() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }
};

// divUp, Common/Commons.h:7-13.
inline int divUp(int a, int b) {
  if (b == 0) return 1;
  return (a % b != 0) ? (a / b + 1) : (a / b);
}

// Number of (i,j) pairs fixBoundaryX/Y/Z reach: the launch is sized for
// total = A*B threads in blocks of min(1024,total) (Fluid/Fluid_3D_common.cu:20-37)
// while the kernel indexes (A+1)*(B+1) pairs (Fluid/Fluid_3D_kernels.cu:46-49).
inline int boundaryCoverage(int A, int B, bool refCoverage) {
  int pairs = (A + 1) * (B + 1);
  if (!refCoverage) return pairs;
  int total = A * B;
  int nt = std::min(1024, total);
  long long launched = (long long)divUp(total, nt) * nt;
  return (int)std::min<long long>(launched, pairs);
}

// Host seeding: Fluid::createParticlesAt / createSquareFluid / createSphereFluid /
// checkCoordValid (Fluid/Fluid_3D_FLIP.cu:508-594), float arithmetic as on the host.
inline bool checkCoordValid(const float c[3]) {  // Fluid_3D_FLIP.cu:542-550
  if (c[0] < 0 || c[1] < 0 || c[2] < 0) return false;
  if (c[0] > 1 || c[1] > 1 || c[2] > 1) return false;
  return true;
}

inline void seedParticles(const aq_flip_config& c, std::vector<float>& pos,
                          std::vector<float>& frac) {
  const int X = c.size[0], Y = c.size[1], Z = c.size[2];
  const float h = c.grid_physical_size / (float)Y;  // Fluid_3D_FLIP.cu:444
  SplitMix64 rng{c.seed};
  pos.clear();
  frac.clear();
  // createParticlesAt, Fluid_3D_FLIP.cu:508-538
  auto createAt = [&](const float cp[3], auto&& filter, float spacing, int phase) {
    for (float dx = 0; dx <= 1; ++dx)
      for (float dy = 0; dy <= 1; ++dy)
        for (float dz = 0; dz <= 1; ++dz) {
          float sub[3] = {(float)(dx - 0.5) * spacing + cp[0], (float)(dy - 0.5) * spacing + cp[1],
                          (float)(dz - 0.5) * spacing + cp[2]};
          float jx = rng.u01() - 0.5f;
          float jy = rng.u01() - 0.5f;
          float jz = rng.u01() - 0.5f;
          float p[3] = {sub[0] + (jx * spacing) * c.particle_jitter,
                        sub[1] + (jy * spacing) * c.particle_jitter,
                        sub[2] + (jz * spacing) * c.particle_jitter};
          float m = spacing / 2;
          const int S[3] = {X, Y, Z};
          for (int a = 0; a < 3; ++a) {
            float hi = (float)S[a] * h - m;
            float v = std::fmin(hi, p[a]);
            p[a] = std::fmax(m, v);
          }
          if (filter(p)) {
            pos.push_back(p[0]);
            pos.push_back(p[1]);
            pos.push_back(p[2]);
            float f[4] = {0, 0, 0, 0};  // Particle(pos, phase), Fluid_3D_FLIP.cuh:31-48
            if (phase >= 0 && phase < 4) f[phase] = 1;
            frac.insert(frac.end(), f, f + 4);
          }
        }
  };
  auto physPos = [&](int x, int y, int z, float out[3]) {  // MAC_Grid_3D.cuh:135-138
    out[0] = ((float)x + 0.5f) * h;
    out[1] = ((float)y + 0.5f) * h;
    out[2] = ((float)z + 0.5f) * h;
  };
  for (int vi = 0; vi < c.n_volumes; ++vi) {
    const aq_flip_volume& v = c.volumes[vi];
    if (v.shape == AQ_SHAPE_BOX) {  // createSquareFluid, Fluid_3D_FLIP.cu:552-569
      if (!checkCoordValid(v.box_min) || !checkCoordValid(v.box_max)) continue;
      for (int z = (int)(v.box_min[2] * (float)Z); (float)z < v.box_max[2] * (float)Z; ++z)
        for (int x = (int)(v.box_min[0] * (float)X); (float)x < v.box_max[0] * (float)X; ++x)
          for (int y = (int)(v.box_min[1] * (float)Y); (float)y < v.box_max[1] * (float)Y; ++y) {
            float cp[3];
            physPos(x, y, z, cp);
            createAt(cp, [](const float*) { return true; }, h / 2, v.phase);
          }
    } else if (v.shape == AQ_SHAPE_SPHERE) {  // createSphereFluid, Fluid_3D_FLIP.cu:571-594
      if (!checkCoordValid(v.center)) continue;
      float cc[3] = {(v.center[0] * (float)X) * h, (v.center[1] * (float)Y) * h,
                     (v.center[2] * (float)Z) * h};
      float rlim = (v.radius * h) * (float)Y;
      auto filter = [&](const float* p) {
        float d0 = p[0] - cc[0], d1 = p[1] - cc[1], d2 = p[2] - cc[2];
        return std::sqrt(d0 * d0 + d1 * d1 + d2 * d2) < rlim;
      };
      for (int y = 0; y < Y; ++y)
        for (int x = 0; x < X; ++x)
          for (int z = 0; z < Z; ++z) {
            float cp[3];
            physPos(x, y, z, cp);
            createAt(cp, filter, h / 2, v.phase);
          }
    }
  }
}

// trilinearUnitKernel / trilinearHatKernel(float3), Fluid/WeightKernels.cuh:131-146.
template <typename T>
inline T hatUnit(T r) {
  r = std::fabs(r);
  if (r > T(1)) return T(0);
  return T(1) - r;
}
template <typename T>
inline T hat3(T rx, T ry, T rz, T support) {
  return hatUnit(rx / support) * hatUnit(ry / support) * hatUnit(rz / support);
}
// helper_math lerp: a + t*(b-a).
template <typename T>
inline T lerp(T a, T b, T t) {
  return a + t * (b - a);
}

template <typename T>
struct Particle {  // struct Particle, Fluid/Fluid_3D_FLIP.cuh:23-49
  T p[3];
  T v[3];
  T f[4];
};

template <typename T>
struct Oracle {
  int X, Y, Z;
  T h;  // cellPhysicalSize = gridPhysicalSize / sizeY (Fluid_3D_FLIP.cu:444)
  aq_flip_config cfg;
  int Np = 0;
  float physicalTime = 0;  // Fluid::physicalTime is a float (Fluid/Fluid.h:11)

  std::vector<Particle<T>> parts, partsCopy;
  std::vector<int> hash, perm, cellStart, cellEnd;
  // Fields of VolumeCollection (Fluid/MAC_Grid_3D.cuh:23-41) with the extents of
  // MAC_Grid_3D's constructor (Fluid/MAC_Grid_3D.cu:53-71).
  std::vector<int> content;         // (N+1)^3
  std::vector<T> vel, nvel;         // (N+1)^3 x 4
  std::vector<T> pressure, ptmp;    // N^3
  std::vector<T> divergence, density;
  std::vector<int> pcount;          // N^3
  std::vector<T> weight;            // N^3 x 4 (velocityAccumWeight)
  std::vector<int> hasvel;          // N^3 x 4
  std::vector<T> vf, nvf;           // N^3 x 4
  std::vector<T> nvelTmp, nvfTmp;
  std::vector<int> hasvelTmp;

  size_t E1() const { return (size_t)(X + 1) * (Y + 1) * (Z + 1); }
  size_t E0() const { return (size_t)X * Y * Z; }
  size_t i1(int x, int y, int z) const { return ((size_t)x * (Y + 1) + y) * (Z + 1) + z; }
  size_t i0(int x, int y, int z) const { return ((size_t)x * Y + y) * Z + z; }

  explicit Oracle(const aq_flip_config& c) : cfg(c) {
    X = c.size[0];
    Y = c.size[1];
    Z = c.size[2];
    h = (T)c.grid_physical_size / (T)Y;
    // MAC_Grid_3D constructor: every field cleared (Fluid/MAC_Grid_3D.cu:53-71).
    content.assign(E1(), AIR);
    vel.assign(E1() * 4, T(0));
    nvel.assign(E1() * 4, T(0));
    pressure.assign(E0(), T(0));
    ptmp.assign(E0(), T(0));
    divergence.assign(E0(), T(0));
    density.assign(E0(), T(0));
    pcount.assign(E0(), 0);
    weight.assign(E0() * 4, T(0));
    hasvel.assign(E0() * 4, 0);
    vf.assign(E0() * 4, T(0));
    nvf.assign(E0() * 4, T(0));
    cellStart.assign(E1(), -1);
    cellEnd.assign(E1(), -1);
  }

  void setParticles(const float* pos, const float* v, const float* f, int n) {
    Np = n;
    parts.resize(n);
    for (int i = 0; i < n; ++i) {
      for (int a = 0; a < 3; ++a) {
        parts[i].p[a] = (T)pos[3 * i + a];
        parts[i].v[a] = v ? (T)v[3 * i + a] : T(0);
      }
      for (int a = 0; a < 4; ++a) parts[i].f[a] = f ? (T)f[4 * i + a] : T(0);
    }
    partsCopy = parts;
    hash.assign(n, 0);
    perm.assign(n, 0);
  }

  // Fluid::init minus rendering (Fluid/Fluid_3D_FLIP.cu:435-505): fixBoundary,
  // then transferToGrid. The mesher's re-sort at :502 is not reproduced.
  void init() {
    fixBoundary();
    transferToGrid();
  }

  // ---- A4-A7: performSpatialHashing2 (Fluid/Fluid_3D_common.cuh:27-56) ----
  void hashAndSort() {
    // writeIndicesImpl (Fluid_3D_kernels.cu:508-514) + calcHashImpl (Fluid_3D_kernels.cuh:8-27)
#pragma omp parallel for
    for (int i = 0; i < Np; ++i) {
      int x = (int)(parts[i].p[0] / h);
      int y = (int)(parts[i].p[1] / h);
      int z = (int)(parts[i].p[2] / h);
      hash[i] = x * (Y * Z) + y * Z + z;
    }
    // sort_by_key(hash, index) (Fluid_3D_common.cuh:35), stable: counting sort in
    // input order.
    int nc = X * Y * Z;
    std::vector<int> off(nc + 1, 0);
    for (int i = 0; i < Np; ++i) off[hash[i] + 1]++;
    for (int c = 0; c < nc; ++c) off[c + 1] += off[c];
    std::vector<int> sortedHash(Np);
    for (int i = 0; i < Np; ++i) {
      int k = off[hash[i]]++;
      perm[k] = i;
      sortedHash[k] = hash[i];
    }
    hash.swap(sortedHash);
    // applySortImpl (Fluid_3D_kernels.cuh:95-102) + std::swap (Fluid_3D_common.cuh:44)
#pragma omp parallel for
    for (int i = 0; i < Np; ++i) partsCopy[i] = parts[perm[i]];
    parts.swap(partsCopy);
    // cudaMemset(0xFF) + findCellStartEndImpl (Fluid_3D_common.cuh:47-48, Fluid_kernels.cuh:3-22)
    std::fill(cellStart.begin(), cellStart.end(), -1);
    std::fill(cellEnd.begin(), cellEnd.end(), -1);
    for (int i = 0; i < Np; ++i) {
      int hh = hash[i];
      if (i == 0 || hash[i - 1] < hh) cellStart[hh] = i;
      if (i == Np - 1 || hash[i + 1] > hh) cellEnd[hh] = i;
    }
  }

  // ---- A8: resetCellImpl (Fluid/Fluid_3D_FLIP.cu:287-302) ----
  void resetCells() {
#pragma omp parallel for
    for (int x = 0; x < X; ++x)
      for (int y = 0; y < Y; ++y)
        for (int z = 0; z < Z; ++z) {
          content[i1(x, y, z)] = AIR;
          for (int a = 0; a < 4; ++a) {
            hasvel[i0(x, y, z) * 4 + a] = 0;
            vel[i1(x, y, z) * 4 + a] = T(0);
            nvel[i1(x, y, z) * 4 + a] = T(0);
          }
        }
  }

  // ---- A9: transferToCellAccumPhase (Fluid/Fluid_3D_FLIP.cu:8-69) ----
  void p2gAccum() {
    const int cellCount = X * Y * Z;
#pragma omp parallel for
    for (int x = 0; x < X; ++x)
      for (int y = 0; y < Y; ++y)
        for (int z = 0; z < Z; ++z) {
          const T xp[3] = {(T)x * h, (T)(y + 0.5) * h, (T)(z + 0.5) * h};
          const T yp[3] = {(T)(x + 0.5) * h, (T)y * h, (T)(z + 0.5) * h};
          const T zp[3] = {(T)(x + 0.5) * h, (T)(y + 0.5) * h, (T)z * h};
          T tv[3] = {0, 0, 0}, w[3] = {0, 0, 0};
          for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
              for (int dz = -1; dz <= 1; ++dz) {
                // Flattened neighbour index, range-checked in 1D only (rows may wrap).
                int cell = (x + dx) * Y * Z + (y + dy) * Z + z + dz;
                if (cell >= 0 && cell < cellCount) {
                  for (int j = cellStart[cell]; j <= cellEnd[cell]; ++j) {
                    if (j >= 0 && j < Np) {
                      const Particle<T>& p = parts[j];
                      T wx = hat3(p.p[0] - xp[0], p.p[1] - xp[1], p.p[2] - xp[2], h);
                      T wy = hat3(p.p[0] - yp[0], p.p[1] - yp[1], p.p[2] - yp[2], h);
                      T wz = hat3(p.p[0] - zp[0], p.p[1] - zp[1], p.p[2] - zp[2], h);
                      tv[0] += wx * p.v[0];
                      tv[1] += wy * p.v[1];
                      tv[2] += wz * p.v[2];
                      w[0] += wx;
                      w[1] += wy;
                      w[2] += wz;
                    }
                  }
                }
              }
          size_t c0 = i0(x, y, z), c1 = i1(x, y, z);
          for (int a = 0; a < 3; ++a) {
            weight[c0 * 4 + a] = w[a];
            vel[c1 * 4 + a] = tv[a];
          }
          weight[c0 * 4 + 3] = T(0);
          vel[c1 * 4 + 3] = T(0);
        }
  }

  // ---- A10: transferToCellDividePhase (Fluid/Fluid_3D_FLIP.cu:72-116) ----
  void p2gDivide() {
#pragma omp parallel for
    for (int x = 0; x < X; ++x)
      for (int y = 0; y < Y; ++y)
        for (int z = 0; z < Z; ++z) {
          size_t c0 = i0(x, y, z), c1 = i1(x, y, z);
          for (int a = 0; a < 3; ++a) {
            T wa = weight[c0 * 4 + a];
            if (wa > 0) {
              vel[c1 * 4 + a] /= wa;
              hasvel[c0 * 4 + a] = 1;
            }
          }
          for (int a = 0; a < 4; ++a) nvel[c1 * 4 + a] = vel[c1 * 4 + a];
          int index = (int)c0;
          for (int j = cellStart[index]; j <= cellEnd[index]; ++j) {
            if (j >= 0 && j < Np) {
              content[c1] = FLUID;
              break;
            }
          }
        }
  }

  // ---- A11: transferVolumeFractionsToCell (Fluid/Fluid_3D_FLIP.cu:118-165) ----
  void p2gFractions() {
    const T* rho = nullptr;
    T pd[4] = {(T)cfg.phase_densities[0], (T)cfg.phase_densities[1], (T)cfg.phase_densities[2],
               (T)cfg.phase_densities[3]};
    rho = pd;
#pragma omp parallel for
    for (int x = 0; x < X; ++x)
      for (int y = 0; y < Y; ++y)
        for (int z = 0; z < Z; ++z) {
          const T cp[3] = {(T)(x + 0.5) * h, (T)(y + 0.5) * h, (T)(z + 0.5) * h};
          T f[4] = {0, 0, 0, 0};
          T wsum = 0;
          int cell = (int)i0(x, y, z);
          for (int j = cellStart[cell]; j <= cellEnd[cell]; ++j) {
            if (j >= 0 && j < Np) {
              const Particle<T>& p = parts[j];
              T w = hat3(p.p[0] - cp[0], p.p[1] - cp[1], p.p[2] - cp[2], h);
              for (int a = 0; a < 4; ++a) f[a] += w * p.f[a];
              wsum += w;
            }
          }
          if (wsum > 0) {
            for (int a = 0; a < 4; ++a) f[a] /= wsum;
            T s = f[0] + f[1] + f[2] + f[3];
            for (int a = 0; a < 4; ++a) f[a] /= s;
          }
          T dens = rho[0] * f[0] + rho[1] * f[1] + rho[2] * f[2] + rho[3] * f[3];
          size_t c0 = (size_t)cell;
          density[c0] = dens;
          for (int a = 0; a < 4; ++a) {
            vf[c0 * 4 + a] = f[a];
            nvf[c0 * 4 + a] = f[a];
          }
        }
  }

  // Fluid::transferToGrid (Fluid/Fluid_3D_FLIP.cu:363-378).
  void transferToGrid() {
    hashAndSort();
    resetCells();
    p2gAccum();
    p2gDivide();
    p2gFractions();
  }

  // ---- A12: applyGravityImpl (Fluid/Fluid_3D_kernels.cu:4-39), gathered per target ----
  void applyGravity() {
    const T dt = (T)cfg.dt;
    const T g[3] = {(T)cfg.gravity[0], (T)cfg.gravity[1], (T)cfg.gravity[2]};
    const T add[3] = {g[0] * dt, g[1] * dt, g[2] * dt};
    auto interior = [&](int x, int y, int z) {
      return x >= 0 && y >= 0 && z >= 0 && x < X && y < Y && z < Z;
    };
    std::vector<T> out(nvel);
#pragma omp parallel for
    for (int x = 0; x <= X; ++x)
      for (int y = 0; y <= Y; ++y)
        for (int z = 0; z <= Z; ++z) {
          size_t c = i1(x, y, z);
          T* o = &out[c * 4];
          int ct = content[c];
          if (interior(x, y, z) && ct == FLUID)
            for (int a = 0; a < 3; ++a) o[a] += add[a];
          if (ct == AIR) {
            const int sx[3] = {x - 1, x, x};
            const int sy[3] = {y, y - 1, y};
            const int sz[3] = {z, z, z - 1};
            for (int d = 0; d < 3; ++d) {
              if (g[d] != 0 && interior(sx[d], sy[d], sz[d]) &&
                  content[i1(sx[d], sy[d], sz[d])] == FLUID)
                for (int a = 0; a < 3; ++a) o[a] += add[a];
            }
          }
        }
    nvel.swap(out);
  }

  // ---- A13: fixBoundaryX/Y/Z (Fluid/Fluid_3D_kernels.cu:44-98; launcher Fluid_3D_common.cu:13-41) ----
  void fixBoundary() {
    bool ref = cfg.ref_boundary_coverage != 0;
    int n = boundaryCoverage(Y, Z, ref);
    for (int i = 0; i < n; ++i) {  // fixBoundaryX
      int y = i / (Z + 1), z = i - y * (Z + 1);
      T& a = nvel[i1(0, y, z) * 4 + 0];
      a = std::fmax(T(0), a);
      T& b = nvel[i1(X, y, z) * 4 + 0];
      b = std::fmin(T(0), b);
      content[i1(X, y, z)] = SOLID;
    }
    n = boundaryCoverage(X, Z, ref);
    for (int i = 0; i < n; ++i) {  // fixBoundaryY
      int x = i / (Z + 1), z = i - x * (Z + 1);
      T& a = nvel[i1(x, 0, z) * 4 + 1];
      a = std::fmax(T(0), a);
      T& b = nvel[i1(x, Y, z) * 4 + 1];
      b = std::fmin(T(0), b);
      content[i1(x, Y, z)] = SOLID;
    }
    n = boundaryCoverage(X, Y, ref);
    for (int i = 0; i < n; ++i) {  // fixBoundaryZ
      int x = i / (Y + 1), y = i - x * (Y + 1);
      T& a = nvel[i1(x, y, 0) * 4 + 2];
      a = std::fmax(T(0), a);
      T& b = nvel[i1(x, y, Z) * 4 + 2];
      b = std::fmin(T(0), b);
      content[i1(x, y, Z)] = SOLID;
    }
  }

  // ---- A14: countParticlesInCellsImpl (Fluid/Fluid_3D_FLIP.cu:168-185) ----
  void countParticles() {
#pragma omp parallel for
    for (int c = 0; c < (int)E0(); ++c) pcount[c] = cellEnd[c] - cellStart[c];
  }

  // ---- A15: computeDivergenceImpl (Fluid/Fluid_3D_kernels.cu:426-458), restPPC = 8 ----
  void computeDivergence() {
    const float restPPC = 8.0f;  // Fluid::particlesPerCell (Fluid_3D_FLIP.cuh:81)
#pragma omp parallel for
    for (int x = 0; x < X; ++x)
      for (int y = 0; y < Y; ++y)
        for (int z = 0; z < Z; ++z) {
          const T* t = &nvel[i1(x, y, z) * 4];
          const T* up = &nvel[i1(x, y + 1, z) * 4];
          const T* rt = &nvel[i1(x + 1, y, z) * 4];
          const T* fr = &nvel[i1(x, y, z + 1) * 4];
          T div = up[1] - t[1] + rt[0] - t[0] + fr[2] - t[2];
          int cnt = pcount[i0(x, y, z)];
          // (currentCount - restParticlesPerCell) is float, * 0.001 is double.
          double corr = (double)((T)cnt - (T)restPPC) * 0.001;
          div = (T)((double)div - corr);
          divergence[i0(x, y, z)] = div;
        }
  }

  // Clamp-to-edge read of an N^3 scalar field.
  T readClamped(const std::vector<T>& f, int x, int y, int z) const {
    x = std::min(std::max(x, 0), X - 1);
    y = std::min(std::max(y, 0), Y - 1);
    z = std::min(std::max(z, 0), Z - 1);
    return f[i0(x, y, z)];
  }

  // ---- A16: jacobiImpl (Fluid/Fluid_3D_kernels.cu:460-505), ping-pong ----
  void jacobiSweep() {
#pragma omp parallel for
    for (int x = 0; x < X; ++x)
      for (int y = 0; y < Y; ++y)
        for (int z = 0; z < Z; ++z) {
          size_t c = i0(x, y, z);
          if (content[i1(x, y, z)] == AIR) {
            ptmp[c] = T(0);
            continue;
          }
          T rhs = -divergence[c] * density[c];
          T np = T(0);
          np += readClamped(pressure, x + 1, y, z);
          np += readClamped(pressure, x - 1, y, z);
          np += readClamped(pressure, x, y + 1, z);
          np += readClamped(pressure, x, y - 1, z);
          np += readClamped(pressure, x, y, z + 1);
          np += readClamped(pressure, x, y, z - 1);
          np += rhs;
          np /= T(6);
          ptmp[c] = np;
        }
    pressure.swap(ptmp);
  }
  // solvePressureJacobi (Fluid/Fluid_3D_common.cu:48-55).
  void solvePressure(int iters) {
    if (cfg.pressure_init == AQ_PRESSURE_ZERO) std::fill(pressure.begin(), pressure.end(), T(0));
    for (int i = 0; i < iters; ++i) jacobiSweep();
  }

  // ---- A17: updateVelocityWithPressureImpl (Fluid/Fluid_3D_kernels.cu:118-207) ----
  void project() {
#pragma omp parallel for
    for (int x = 0; x < X; ++x)
      for (int y = 0; y < Y; ++y)
        for (int z = 0; z < Z; ++z) {
          int hv[3] = {0, 0, 0};
          int tc = content[i1(x, y, z)];
          T tp = pressure[i0(x, y, z)];
          T td = density[i0(x, y, z)];
          T* nv = &nvel[i1(x, y, z) * 4];
          const int nx[3] = {x - 1, x, x}, ny[3] = {y, y - 1, y}, nz[3] = {z, z, z - 1};
          const int coord[3] = {x, y, z};
          for (int d = 0; d < 3; ++d) {
            if (coord[d] > 0) {
              int oc = content[i1(nx[d], ny[d], nz[d])];
              T op = pressure[i0(nx[d], ny[d], nz[d])];
              if (tc == FLUID || oc == FLUID) {
                T used = td;
                T od = density[i0(nx[d], ny[d], nz[d])];
                if (tc == FLUID && oc == FLUID)
                  used = (od + td) / T(2);
                else if (oc == FLUID)
                  used = od;
                nv[d] = nv[d] - (tp - op) / used;
                hv[d] = 1;
              }
            }
          }
          int* H = &hasvel[i0(x, y, z) * 4];
          H[0] = hv[0];
          H[1] = hv[1];
          H[2] = hv[2];
          H[3] = 0;
        }
  }

  // ---- A18: extrapolateVelocityByOne (Fluid/Fluid_3D_kernels.cu:210-422), ping-pong ----
  void extrapolateOnce() {
    const T eps = (T)(-1 + 1e-6);
    nvelTmp = nvel;  // faces outside the launch domain are carried over unchanged
    hasvelTmp = hasvel;
#pragma omp parallel for
    for (int x = 0; x < X; ++x)
      for (int y = 0; y < Y; ++y)
        for (int z = 0; z < Z; ++z) {
          int H[3] = {hasvel[i0(x, y, z) * 4 + 0], hasvel[i0(x, y, z) * 4 + 1],
                      hasvel[i0(x, y, z) * 4 + 2]};
          T V[3] = {nvel[i1(x, y, z) * 4 + 0], nvel[i1(x, y, z) * 4 + 1],
                    nvel[i1(x, y, z) * 4 + 2]};
          for (int comp = 0; comp < 3; ++comp) {
            if (H[comp]) continue;
            T sum = 0;
            int cnt = 0;
            // neighbour order: -x, -y, -z, +x, +y, +z; a neighbour in direction d
            // is used if its own velocity along d points away from this face's cell.
            struct N {
              int ok, x, y, z, d, sgn;
            } nb[6] = {{x > 0, x - 1, y, z, 0, +1},     {y > 0, x, y - 1, z, 1, +1},
                       {z > 0, x, y, z - 1, 2, +1},     {x < X - 1, x + 1, y, z, 0, -1},
                       {y < Y - 1, x, y + 1, z, 1, -1}, {z < Z - 1, x, y, z + 1, 2, -1}};
            for (int k = 0; k < 6; ++k) {
              if (!nb[k].ok) continue;
              int nh = hasvel[i0(nb[k].x, nb[k].y, nb[k].z) * 4 + comp];
              const T* nv = &nvel[i1(nb[k].x, nb[k].y, nb[k].z) * 4];
              bool dirOk = nb[k].sgn > 0 ? (nv[nb[k].d] > eps) : (nv[nb[k].d] < -eps);
              if (nh && dirOk) {
                sum += nv[comp];
                cnt++;
              }
            }
            if (cnt > 0) {
              V[comp] = sum / (T)cnt;
              H[comp] = 1;
            }
          }
          for (int a = 0; a < 3; ++a) {
            nvelTmp[i1(x, y, z) * 4 + a] = V[a];
            hasvelTmp[i0(x, y, z) * 4 + a] = H[a];
          }
        }
    nvel.swap(nvelTmp);
    hasvel.swap(hasvelTmp);
  }
  // extrapolateVelocity: maxDistance = 3 (Fluid/Fluid_3D_common.cu:65-82).
  void extrapolate() {
    for (int i = 0; i < 3; ++i) extrapolateOnce();
  }

  // ---- A19: diffusionJacobiImpl (Fluid/Fluid_3D_kernels.cu:520-556), ping-pong ----
  void diffusionSweep(T lambda) {
    nvfTmp = nvf;
#pragma omp parallel for
    for (int x = 0; x < X; ++x)
      for (int y = 0; y < Y; ++y)
        for (int z = 0; z < Z; ++z) {
          size_t c = i0(x, y, z);
          if (content[i1(x, y, z)] == AIR) {
            pressure[c] = T(0);  // the reference zeroes pressure here
            continue;
          }
          T r[4] = {0, 0, 0, 0};
          const T cc = T(1) + T(2) * lambda;
          const int nbx[6] = {x + 1, x - 1, x, x, x, x};
          const int nby[6] = {y, y, y + 1, y - 1, y, y};
          const int nbz[6] = {z, z, z, z, z + 1, z - 1};
          for (int k = 0; k < 6; ++k) {
            int xx = std::min(std::max(nbx[k], 0), X - 1);
            int yy = std::min(std::max(nby[k], 0), Y - 1);
            int zz = std::min(std::max(nbz[k], 0), Z - 1);
            const T* nf = &nvf[i0(xx, yy, zz) * 4];
            for (int a = 0; a < 4; ++a) r[a] += nf[a] * lambda;
          }
          for (int a = 0; a < 4; ++a) r[a] += nvf[c * 4 + a];
          for (int a = 0; a < 4; ++a) r[a] /= cc;
          T s = r[0] + r[1] + r[2] + r[3];
          for (int a = 0; a < 4; ++a) r[a] /= s;
          for (int a = 0; a < 4; ++a) nvfTmp[c * 4 + a] = r[a];
        }
    nvf.swap(nvfTmp);
  }
  // solveDiffusionJacobi, lambda = D*dt/dx^2 (Fluid/Fluid_3D_common.cu:85-91).
  void diffusion(int iters) {
    T lambda = ((T)cfg.diffusion_coeff * (T)cfg.dt) / (h * h);
    for (int i = 0; i < iters; ++i) diffusionSweep(lambda);
  }

  // ---- getInterpolatedValueAtPoint (Fluid/MAC_Grid_3D.cuh:66-97), one component ----
  T interp(const std::vector<T>& field, int comp, T x, T y, T z) const {
    x = std::fmax(std::fmin(x, (T)X - T(1)), T(0));
    y = std::fmax(std::fmin(y, (T)Y - T(1)), T(0));
    z = std::fmax(std::fmin(z, (T)Z - T(1)), T(0));
    int i = (int)std::floor(x), j = (int)std::floor(y), k = (int)std::floor(z);
    T tx = x - (T)i, ty = y - (T)j, tz = z - (T)k;
    auto at = [&](int a, int b, int c) { return field[i1(a, b, c) * 4 + comp]; };
    T x0y0 = lerp(at(i, j, k), at(i, j, k + 1), tz);
    T x0y1 = lerp(at(i, j + 1, k), at(i, j + 1, k + 1), tz);
    T x1y0 = lerp(at(i + 1, j, k), at(i + 1, j, k + 1), tz);
    T x1y1 = lerp(at(i + 1, j + 1, k), at(i + 1, j + 1, k + 1), tz);
    T x0 = lerp(x0y0, x0y1, ty);
    T x1 = lerp(x1y0, x1y1, ty);
    return lerp(x0, x1, tx);
  }
  // getPointVelocity / getPointNewVelocity (Fluid/MAC_Grid_3D.cuh:100-130).
  void pointVelocity(const std::vector<T>& field, const T p[3], T out[3]) const {
    T x = p[0] / h, y = p[1] / h, z = p[2] / h;
    out[0] = interp(field, 0, x, (T)(y - T(0.5)), (T)(z - T(0.5)));
    out[1] = interp(field, 1, (T)(x - T(0.5)), y, (T)(z - T(0.5)));
    out[2] = interp(field, 2, (T)(x - T(0.5)), (T)(y - T(0.5)), z);
  }

  // ---- A20: transferToParticlesImpl (Fluid/Fluid_3D_FLIP.cu:189-227) ----
  void g2p() {
    const T a = (T)cfg.flip_coeff;
#pragma omp parallel for
    for (int i = 0; i < Np; ++i) {
      Particle<T>& P = parts[i];
      T nv[3], ov[3];
      pointVelocity(nvel, P.p, nv);
      pointVelocity(vel, P.p, ov);
      for (int c = 0; c < 3; ++c) {
        T change = nv[c] - ov[c];
        T flip = P.v[c] + change;
        T pic = nv[c];
        P.v[c] = a * flip + (T(1) - a) * pic;
      }
      int cx = (int)std::floor(P.p[0] / h);
      int cy = (int)std::floor(P.p[1] / h);
      int cz = (int)std::floor(P.p[2] / h);
      const T* f = &nvf[i0(cx, cy, cz) * 4];
      for (int c = 0; c < 4; ++c) P.f[c] = f[c];
      T s = P.f[0] * T(1) + P.f[1] * T(1) + P.f[2] * T(1) + P.f[3] * T(1);
      if (s > 0)
        for (int c = 0; c < 4; ++c) P.f[c] /= s;
    }
  }

  // ---- A21: moveParticlesImpl (Fluid/Fluid_3D_FLIP.cu:229-285), RK3 ----
  void advect() {
    const T dt = (T)cfg.dt;
    const T m = h / T(4);
    const T L[3] = {(T)X * h, (T)Y * h, (T)Z * h};
#pragma omp parallel for
    for (int i = 0; i < Np; ++i) {
      Particle<T>& P = parts[i];
      T b[3] = {P.p[0], P.p[1], P.p[2]};
      T u1[3], u2[3], u3[3], q[3];
      pointVelocity(nvel, b, u1);
      for (int c = 0; c < 3; ++c) q[c] = b[c] + (dt * u1[c]) / T(2);
      pointVelocity(nvel, q, u2);
      for (int c = 0; c < 3; ++c) q[c] = b[c] + ((dt * u2[c]) * T(3)) / T(4);
      pointVelocity(nvel, q, u3);
      T d[3];
      for (int c = 0; c < 3; ++c)
        d[c] = b[c] + dt * ((((u1[c] * T(2)) / T(9)) + ((u2[c] * T(3)) / T(9))) + ((u3[c] * T(4)) / T(9)));
      for (int c = 0; c < 3; ++c) {
        if (d[c] < m) {
          d[c] = m;
          P.v[c] *= T(-1);
        }
        if (d[c] > L[c] - m) {
          d[c] = L[c] - m;
          P.v[c] *= T(-1);
        }
      }
      for (int c = 0; c < 3; ++c) P.p[c] = d[c];
    }
  }

  // ---- A24: Fluid::simulationStep (Fluid/Fluid_3D_FLIP.cu:316-352) ----
  void step() {
    g2p();
    advect();
    transferToGrid();
    applyGravity();
    fixBoundary();
    countParticles();
    computeDivergence();
    solvePressure(cfg.pressure_iters);
    project();
    extrapolate();
    diffusion(cfg.diffusion_iters);
    physicalTime += cfg.dt;
  }

  // The same step with the wall-clock seconds of each of its stages added to out[9]:
  // g2p, advect, transfer to grid (hash, sort, P2G), gravity + boundary, count +
  // divergence, pressure sweeps, projection, extrapolation, diffusion sweeps.
  template <typename Clock>
  void stepTimed(double out[9]) {
    auto t = Clock::now();
    auto lap = [&](int k) {
      const auto n = Clock::now();
      out[k] += std::chrono::duration<double>(n - t).count();
      t = n;
    };
    g2p();
    lap(0);
    advect();
    lap(1);
    transferToGrid();
    lap(2);
    applyGravity();
    fixBoundary();
    lap(3);
    countParticles();
    computeDivergence();
    lap(4);
    solvePressure(cfg.pressure_iters);
    lap(5);
    project();
    lap(6);
    extrapolate();
    lap(7);
    diffusion(cfg.diffusion_iters);
    lap(8);
    physicalTime += cfg.dt;
  }

  // ---- N1: FP64 relative residual of the pressure system ----
  // A: diagonal 6 - #(out-of-domain neighbours), -1 per FLUID neighbour, AIR = 0
  // (Fluid/Legacy/PressureEquation.cuh:216-328); b = -div*rho as jacobiImpl forms it.
  double residual() const {
    double rr = 0, bb = 0;
    for (int x = 0; x < X; ++x)
      for (int y = 0; y < Y; ++y)
        for (int z = 0; z < Z; ++z) {
          if (content[i1(x, y, z)] != FLUID) continue;
          size_t c = i0(x, y, z);
          double b = (double)(T)(-divergence[c] * density[c]);
          double pc = (double)pressure[c];
          double ap = 6.0 * pc;
          const int nx[6] = {x + 1, x - 1, x, x, x, x}, ny[6] = {y, y, y + 1, y - 1, y, y},
                    nz[6] = {z, z, z, z, z + 1, z - 1};
          for (int k = 0; k < 6; ++k) {
            if (nx[k] < 0 || ny[k] < 0 || nz[k] < 0 || nx[k] >= X || ny[k] >= Y || nz[k] >= Z) {
              ap -= pc;
            } else if (content[i1(nx[k], ny[k], nz[k])] == FLUID) {
              ap -= (double)pressure[i0(nx[k], ny[k], nz[k])];
            }
          }
          double r = b - ap;
          rr += r * r;
          bb += b * b;
        }
    if (bb == 0) return std::sqrt(rr);
    return std::sqrt(rr) / std::sqrt(bb);
  }
};

}  // namespace aqo
