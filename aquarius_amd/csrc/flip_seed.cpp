// flip_seed.cpp -- host particle seeding of libaqflip.
//
// Fluid::createParticlesAt / createSquareFluid / createSphereFluid /
// checkCoordValid (Fluid/Fluid_3D_FLIP.cu:508-594): per grid cell a 2x2x2 lattice
// of sub-particles at spacing dx/2, each jittered by (U - 0.5)*spacing*jitter,
// clamped to [spacing/2, L - spacing/2] and kept if it passes the volume's filter.
// Boxes visit cells z, x, y; spheres visit every cell y, x, z. The jitter stream is
// a seeded SplitMix64 (24-bit uniform floats in [0,1)) in place of the reference's
// unseeded rand() (Common/Commons.h:3-5); three draws per sub-particle in x, y, z
// order, drawn before the filter as in the reference, so the stream position of
// every particle is fixed by the scene.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/aqflip.h"

namespace aqf {

namespace {

struct JitterStream {
  uint64_t state;
  float next01() {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (float)(z >> 40) * (1.0f / 16777216.0f);
  }
};

bool inUnitCube(const float c[3]) {
  for (int a = 0; a < 3; ++a)
    if (c[a] < 0 || c[a] > 1) return false;
  return true;
}

struct Seeder {
  const aq_flip_config& cfg;
  float h;
  JitterStream rng;
  std::vector<float>* pos;
  std::vector<float>* frac;
  int64_t count = 0;

  // Emit the 8 jittered sub-particles around one cell centre.
  template <typename Keep>
  void cell(const float centre[3], float spacing, int phase, Keep keep) {
    const int S[3] = {cfg.size[0], cfg.size[1], cfg.size[2]};
    const float half = spacing / 2;
    for (int o = 0; o < 8; ++o) {
      const float lattice[3] = {(o & 4) ? 0.5f : -0.5f, (o & 2) ? 0.5f : -0.5f,
                                (o & 1) ? 0.5f : -0.5f};
      float q[3];
      float jit[3];
      for (int a = 0; a < 3; ++a) jit[a] = rng.next01() - 0.5f;
      for (int a = 0; a < 3; ++a) {
        const float sub = lattice[a] * spacing + centre[a];
        float v = sub + (jit[a] * spacing) * cfg.particle_jitter;
        const float upper = (float)S[a] * h - half;
        v = std::fmin(upper, v);
        q[a] = std::fmax(half, v);
      }
      if (!keep(q)) continue;
      ++count;
      if (pos) {
        pos->insert(pos->end(), q, q + 3);
        float f[4] = {0.f, 0.f, 0.f, 0.f};
        if (phase >= 0 && phase < 4) f[phase] = 1.f;
        frac->insert(frac->end(), f, f + 4);
      }
    }
  }

  void centreOf(int x, int y, int z, float c[3]) const {
    c[0] = ((float)x + 0.5f) * h;
    c[1] = ((float)y + 0.5f) * h;
    c[2] = ((float)z + 0.5f) * h;
  }

  void box(const aq_flip_volume& v) {
    if (!inUnitCube(v.box_min) || !inUnitCube(v.box_max)) return;
    const int X = cfg.size[0], Y = cfg.size[1], Z = cfg.size[2];
    for (int z = (int)(v.box_min[2] * (float)Z); (float)z < v.box_max[2] * (float)Z; ++z)
      for (int x = (int)(v.box_min[0] * (float)X); (float)x < v.box_max[0] * (float)X; ++x)
        for (int y = (int)(v.box_min[1] * (float)Y); (float)y < v.box_max[1] * (float)Y; ++y) {
          float c[3];
          centreOf(x, y, z, c);
          cell(c, h / 2, v.phase, [](const float*) { return true; });
        }
  }

  void sphere(const aq_flip_volume& v) {
    if (!inUnitCube(v.center)) return;
    const int X = cfg.size[0], Y = cfg.size[1], Z = cfg.size[2];
    const float cc[3] = {(v.center[0] * (float)X) * h, (v.center[1] * (float)Y) * h,
                         (v.center[2] * (float)Z) * h};
    const float limit = (v.radius * h) * (float)Y;
    auto inside = [&](const float* q) {
      const float d0 = q[0] - cc[0], d1 = q[1] - cc[1], d2 = q[2] - cc[2];
      return std::sqrt(d0 * d0 + d1 * d1 + d2 * d2) < limit;
    };
    for (int y = 0; y < Y; ++y)
      for (int x = 0; x < X; ++x)
        for (int z = 0; z < Z; ++z) {
          float c[3];
          centreOf(x, y, z, c);
          cell(c, h / 2, v.phase, inside);
        }
  }
};

}  // namespace

// Fills pos (n*3) and frac (n*4) when non-null; returns the particle count.
int64_t seed_scene(const aq_flip_config& cfg, std::vector<float>* pos, std::vector<float>* frac) {
  Seeder s{cfg, cfg.grid_physical_size / (float)cfg.size[1], JitterStream{cfg.seed}, pos, frac};
  if (pos) {
    pos->clear();
    frac->clear();
  }
  for (int i = 0; i < cfg.n_volumes; ++i) {
    const aq_flip_volume& v = cfg.volumes[i];
    if (v.shape == AQ_SHAPE_BOX)
      s.box(v);
    else if (v.shape == AQ_SHAPE_SPHERE)
      s.sphere(v);
  }
  return s.count;
}

}  // namespace aqf
