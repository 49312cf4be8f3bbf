// flip_oracle.cpp -- C entry points of the CPU oracle (TEST INFRASTRUCTURE ONLY).
//
// Exposes aqo::Oracle<float> and aqo::Oracle<double> to the ctypes test harness,
// the smoke check and bench.py's CPU-baseline leg. Fields are exchanged in the
// same host layout the GPU library exports (include/aqflip.h, AQ_FIELD_*).
#include <omp.h>

#include <chrono>
#include <cstdio>

#include "flip_oracle.hpp"

namespace {

struct Handle {
  int fp64;
  aqo::Oracle<float>* f = nullptr;
  aqo::Oracle<double>* d = nullptr;
};

template <typename O, typename Fn>
void dispatch(Handle* h, Fn&& fn) {
  if (h->fp64)
    fn(*h->d);
  else
    fn(*h->f);
}

// Element count and width of each exported field.
template <typename T>
void fieldShape(const aqo::Oracle<T>& o, int field, size_t* n, int* width) {
  switch (field) {
    case AQ_FIELD_CONTENT: *n = o.E1(); *width = 1; return;
    case AQ_FIELD_VELOCITY:
    case AQ_FIELD_NEW_VELOCITY: *n = o.E1(); *width = 4; return;
    case AQ_FIELD_PRESSURE:
    case AQ_FIELD_DIVERGENCE:
    case AQ_FIELD_DENSITY:
    case AQ_FIELD_PARTICLE_COUNT: *n = o.E0(); *width = 1; return;
    case AQ_FIELD_HAS_VELOCITY:
    case AQ_FIELD_VOLUME_FRACTIONS:
    case AQ_FIELD_NEW_VOLUME_FRACTIONS: *n = o.E0(); *width = 4; return;
    default: *n = 0; *width = 0;
  }
}

template <typename T>
void getField(aqo::Oracle<T>& o, int field, double* dst) {
  size_t n;
  int w;
  fieldShape(o, field, &n, &w);
  size_t m = n * w;
  switch (field) {
    case AQ_FIELD_CONTENT: for (size_t i = 0; i < m; ++i) dst[i] = o.content[i]; break;
    case AQ_FIELD_VELOCITY: for (size_t i = 0; i < m; ++i) dst[i] = o.vel[i]; break;
    case AQ_FIELD_NEW_VELOCITY: for (size_t i = 0; i < m; ++i) dst[i] = o.nvel[i]; break;
    case AQ_FIELD_PRESSURE: for (size_t i = 0; i < m; ++i) dst[i] = o.pressure[i]; break;
    case AQ_FIELD_DIVERGENCE: for (size_t i = 0; i < m; ++i) dst[i] = o.divergence[i]; break;
    case AQ_FIELD_DENSITY: for (size_t i = 0; i < m; ++i) dst[i] = o.density[i]; break;
    case AQ_FIELD_PARTICLE_COUNT: for (size_t i = 0; i < m; ++i) dst[i] = o.pcount[i]; break;
    case AQ_FIELD_HAS_VELOCITY: for (size_t i = 0; i < m; ++i) dst[i] = o.hasvel[i]; break;
    case AQ_FIELD_VOLUME_FRACTIONS: for (size_t i = 0; i < m; ++i) dst[i] = o.vf[i]; break;
    case AQ_FIELD_NEW_VOLUME_FRACTIONS: for (size_t i = 0; i < m; ++i) dst[i] = o.nvf[i]; break;
  }
}

template <typename T>
void setField(aqo::Oracle<T>& o, int field, const double* src) {
  size_t n;
  int w;
  fieldShape(o, field, &n, &w);
  size_t m = n * w;
  switch (field) {
    case AQ_FIELD_CONTENT: for (size_t i = 0; i < m; ++i) o.content[i] = (int)src[i]; break;
    case AQ_FIELD_VELOCITY: for (size_t i = 0; i < m; ++i) o.vel[i] = (T)src[i]; break;
    case AQ_FIELD_NEW_VELOCITY: for (size_t i = 0; i < m; ++i) o.nvel[i] = (T)src[i]; break;
    case AQ_FIELD_PRESSURE: for (size_t i = 0; i < m; ++i) o.pressure[i] = (T)src[i]; break;
    case AQ_FIELD_DIVERGENCE: for (size_t i = 0; i < m; ++i) o.divergence[i] = (T)src[i]; break;
    case AQ_FIELD_DENSITY: for (size_t i = 0; i < m; ++i) o.density[i] = (T)src[i]; break;
    case AQ_FIELD_PARTICLE_COUNT: for (size_t i = 0; i < m; ++i) o.pcount[i] = (int)src[i]; break;
    case AQ_FIELD_HAS_VELOCITY: for (size_t i = 0; i < m; ++i) o.hasvel[i] = (int)src[i]; break;
    case AQ_FIELD_VOLUME_FRACTIONS: for (size_t i = 0; i < m; ++i) o.vf[i] = (T)src[i]; break;
    case AQ_FIELD_NEW_VOLUME_FRACTIONS: for (size_t i = 0; i < m; ++i) o.nvf[i] = (T)src[i]; break;
  }
}

template <typename T>
void runStage(aqo::Oracle<T>& o, int stage) {
  switch (stage) {
    case AQ_STAGE_G2P: o.g2p(); break;
    case AQ_STAGE_ADVECT: o.advect(); break;
    case AQ_STAGE_TRANSFER_TO_GRID: o.transferToGrid(); break;
    case AQ_STAGE_GRAVITY: o.applyGravity(); break;
    case AQ_STAGE_BOUNDARY: o.fixBoundary(); break;
    case AQ_STAGE_DIVERGENCE: o.countParticles(); o.computeDivergence(); break;
    case AQ_STAGE_PRESSURE: o.solvePressure(o.cfg.pressure_iters); break;
    case AQ_STAGE_PROJECT: o.project(); break;
    case AQ_STAGE_EXTRAPOLATE: o.extrapolate(); break;
    case AQ_STAGE_DIFFUSION: o.diffusion(o.cfg.diffusion_iters); break;
  }
}

}  // namespace

extern "C" {

void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_max_threads(void) { return omp_get_max_threads(); }

// Host seeding restated from Fluid/Fluid_3D_FLIP.cu:508-594 (float arithmetic).
int orc_seed(const aq_flip_config* cfg, float* pos, float* frac4, int64_t cap, int64_t* n) {
  std::vector<float> p, f;
  aqo::seedParticles(*cfg, p, f);
  int64_t cnt = (int64_t)(p.size() / 3);
  *n = cnt;
  if (!pos) return 0;
  if (cap < cnt) return 1;
  std::memcpy(pos, p.data(), p.size() * sizeof(float));
  if (frac4) std::memcpy(frac4, f.data(), f.size() * sizeof(float));
  return 0;
}

void* orc_create(const aq_flip_config* cfg, int fp64) {
  Handle* h = new Handle;
  h->fp64 = fp64;
  if (fp64)
    h->d = new aqo::Oracle<double>(*cfg);
  else
    h->f = new aqo::Oracle<float>(*cfg);
  return h;
}

void orc_destroy(void* p) {
  Handle* h = (Handle*)p;
  delete h->f;
  delete h->d;
  delete h;
}

void orc_set_particles(void* p, const float* pos, const float* vel, const float* frac4, int n) {
  dispatch<void>((Handle*)p, [&](auto& o) { o.setParticles(pos, vel, frac4, n); });
}

int orc_particle_count(void* p) {
  int n = 0;
  dispatch<void>((Handle*)p, [&](auto& o) { n = o.Np; });
  return n;
}

void orc_get_particles(void* p, double* pos, double* vel, double* frac4) {
  dispatch<void>((Handle*)p, [&](auto& o) {
    for (int i = 0; i < o.Np; ++i) {
      for (int a = 0; a < 3; ++a) {
        if (pos) pos[3 * i + a] = o.parts[i].p[a];
        if (vel) vel[3 * i + a] = o.parts[i].v[a];
      }
      if (frac4)
        for (int a = 0; a < 4; ++a) frac4[4 * i + a] = o.parts[i].f[a];
    }
  });
}

void orc_init(void* p) {
  dispatch<void>((Handle*)p, [&](auto& o) { o.init(); });
}
void orc_step(void* p) {
  dispatch<void>((Handle*)p, [&](auto& o) { o.step(); });
}
void orc_stage(void* p, int stage) {
  dispatch<void>((Handle*)p, [&](auto& o) { runStage(o, stage); });
}
// Finer-grained pieces of transferToGrid, for per-kernel checks.
//   0 hashAndSort, 1 resetCells, 2 p2gAccum, 3 p2gDivide, 4 p2gFractions,
//   5 countParticles, 6 computeDivergence, 7 one Jacobi sweep (no zeroing),
//   8 one extrapolation pass, 9 one diffusion sweep.
void orc_substage(void* p, int id) {
  dispatch<void>((Handle*)p, [&](auto& o) {
    using T = typename std::remove_reference<decltype(o)>::type;
    (void)sizeof(T);
    switch (id) {
      case 0: o.hashAndSort(); break;
      case 1: o.resetCells(); break;
      case 2: o.p2gAccum(); break;
      case 3: o.p2gDivide(); break;
      case 4: o.p2gFractions(); break;
      case 5: o.countParticles(); break;
      case 6: o.computeDivergence(); break;
      case 7: o.jacobiSweep(); break;
      case 8: o.extrapolateOnce(); break;
      case 9: o.diffusion(1); break;
    }
  });
}
void orc_jacobi_sweeps(void* p, int n) {
  dispatch<void>((Handle*)p, [&](auto& o) {
    for (int i = 0; i < n; ++i) o.jacobiSweep();
  });
}

size_t orc_field_elems(void* p, int field) {
  size_t n = 0;
  int w = 0;
  dispatch<void>((Handle*)p, [&](auto& o) { fieldShape(o, field, &n, &w); });
  return n * (size_t)w;
}
void orc_get_field(void* p, int field, double* dst) {
  dispatch<void>((Handle*)p, [&](auto& o) { getField(o, field, dst); });
}
void orc_set_field(void* p, int field, const double* src) {
  dispatch<void>((Handle*)p, [&](auto& o) { setField(o, field, src); });
}

void orc_get_cells(void* p, int* start, int* end, int* hash, int* perm) {
  dispatch<void>((Handle*)p, [&](auto& o) {
    if (start) std::memcpy(start, o.cellStart.data(), o.cellStart.size() * sizeof(int));
    if (end) std::memcpy(end, o.cellEnd.data(), o.cellEnd.size() * sizeof(int));
    if (hash) std::memcpy(hash, o.hash.data(), o.hash.size() * sizeof(int));
    if (perm) std::memcpy(perm, o.perm.data(), o.perm.size() * sizeof(int));
  });
}

double orc_residual(void* p) {
  double r = 0;
  dispatch<void>((Handle*)p, [&](auto& o) { r = o.residual(); });
  return r;
}

double orc_physical_time(void* p) {
  double t = 0;
  dispatch<void>((Handle*)p, [&](auto& o) { t = o.physicalTime; });
  return t;
}

// Wall-clock seconds of `steps` full steps (CPU baseline timing).
// One step timed stage by stage (steady_clock): adds each stage's seconds to
// stage_seconds[9] (see Oracle::stepTimed) and returns the step's total.
double orc_time_step_stages(void* p, double* stage_seconds) {
  double local[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  auto t0 = std::chrono::steady_clock::now();
  dispatch<void>((Handle*)p, [&](auto& o) { o.template stepTimed<std::chrono::steady_clock>(local); });
  auto t1 = std::chrono::steady_clock::now();
  for (int k = 0; k < 9; ++k) stage_seconds[k] += local[k];
  return std::chrono::duration<double>(t1 - t0).count();
}

double orc_time_steps(void* p, int steps) {
  auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < steps; ++i) orc_step(p);
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
