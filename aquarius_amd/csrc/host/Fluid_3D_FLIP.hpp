// Fluid_3D_FLIP.hpp -- drop-in C++ host surface of the reference solver over the
// libaqflip C ABI (include/aqflip.h).
//
// Mirrors, with the same names, members and defaults:
//   ShapeType, InitializationVolume, ConfigFLIP, FluidConfig  Fluid/FluidConfig.cuh:16-97
//   class Fluid, class Fluid_3D                              Fluid/Fluid.h:4-12, Fluid/Fluid_3D.cuh:7-13
//   Fluid_3D_FLIP::Fluid (init / simulationStep / draw /
//   getCenter / getContainerSize / physicalTime)             Fluid/Fluid_3D_FLIP.cuh:53-156
// Differences a caller sees: getCenter returns aqhost::float3 instead of glm::vec3,
// draw() does nothing (rendering is not part of this library), errors throw
// std::runtime_error carrying aq_last_error() instead of exiting the process, and
// FluidConfig carries four extra members (pressureInit, refBoundaryCoverage, seed,
// gridPhysicalSize) whose defaults reproduce the reference's behaviour.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "aqflip.h"

namespace aqhost {
struct float3 {
  float x, y, z;
};
struct float4 {
  float x, y, z, w;
};
inline float3 make_float3(float x, float y, float z) { return {x, y, z}; }
inline float4 make_float4(float x, float y, float z, float w) { return {x, y, z, w}; }
}  // namespace aqhost

using aqhost::float3;
using aqhost::float4;
using aqhost::make_float3;
using aqhost::make_float4;

enum class ShapeType { Sphere, Square };

struct InitializationVolume {
  ShapeType shapeType = ShapeType::Square;
  float3 boxMin = make_float3(0, 0, 0);
  float3 boxMax = make_float3(1, 0.1875f, 1);
  float3 ballCenter = make_float3(0.5f, 0.75f, 0.5f);
  float ballRadius = 0.125f;
  int phase = 0;
  InitializationVolume() {}
  InitializationVolume(ShapeType type) { shapeType = type; }
};

struct ConfigFLIP {
  int sizeX = 50;
  int sizeY = 50;
  int sizeZ = 50;
  int pressureIterations = 100;
  int diffusionIterations = 100;
  float timestep = 0.033f;
  float FLIPcoeff = 0.95f;
  float particleJitter = 0.05f;
};

struct FluidConfig {
  std::string method = "FLIP";
  std::vector<InitializationVolume> initialVolumes;
  float3 gravity = make_float3(0, -9.8f, 0);
  int phaseCount = 2;
  std::vector<float4> phaseColors;
  float4 phaseDensities = make_float4(1, 1, 1, 1);
  float diffusionCoeff = 0.001f;
  ConfigFLIP FLIP;
  // Added by this library (defaults = reference behaviour).
  int pressureInit = AQ_PRESSURE_PREVIOUS;
  int refBoundaryCoverage = 1;
  unsigned long long seed = 0;
  float gridPhysicalSize = 10.f;

  FluidConfig() {
    initialVolumes.emplace_back(ShapeType::Square);
    initialVolumes.emplace_back(ShapeType::Sphere);
    phaseColors.push_back(make_float4(0, 0, 1, 0.2f));
    phaseColors.push_back(make_float4(1, 0, 0, 1));
  }
};

struct DrawCommand {};  // rendering is not part of this library

class Fluid {
 public:
  virtual void simulationStep() = 0;
  virtual void draw(const DrawCommand& drawCommand) = 0;
  virtual void init(FluidConfig config) = 0;
  virtual ~Fluid() {}
  float physicalTime = 0;  // Time elapsed in the simulation.
};

class Fluid_3D : public Fluid {
 public:
  virtual float3 getCenter() = 0;
  virtual float getContainerSize() = 0;
  virtual ~Fluid_3D() {}
};

namespace Fluid_3D_FLIP {

inline void check(aq_status s) {
  if (s != AQ_OK) throw std::runtime_error(std::string("aqflip: ") + aq_last_error());
}

// FluidConfig -> aq_flip_config.
inline aq_flip_config toAbi(const FluidConfig& c) {
  aq_flip_config a;
  check(aq_flip_config_default(&a));
  a.size[0] = c.FLIP.sizeX;
  a.size[1] = c.FLIP.sizeY;
  a.size[2] = c.FLIP.sizeZ;
  a.dt = c.FLIP.timestep;
  a.flip_coeff = c.FLIP.FLIPcoeff;
  a.particle_jitter = c.FLIP.particleJitter;
  a.grid_physical_size = c.gridPhysicalSize;
  a.pressure_iters = c.FLIP.pressureIterations;
  a.diffusion_iters = c.FLIP.diffusionIterations;
  a.phase_count = c.phaseCount;
  a.gravity[0] = c.gravity.x;
  a.gravity[1] = c.gravity.y;
  a.gravity[2] = c.gravity.z;
  a.phase_densities[0] = c.phaseDensities.x;
  a.phase_densities[1] = c.phaseDensities.y;
  a.phase_densities[2] = c.phaseDensities.z;
  a.phase_densities[3] = c.phaseDensities.w;
  a.diffusion_coeff = c.diffusionCoeff;
  a.pressure_init = c.pressureInit;
  a.ref_boundary_coverage = c.refBoundaryCoverage;
  a.seed = c.seed;
  if (c.initialVolumes.size() > AQ_MAX_VOLUMES) throw std::runtime_error("too many initial volumes");
  a.n_volumes = (int)c.initialVolumes.size();
  for (size_t i = 0; i < c.initialVolumes.size(); ++i) {
    const InitializationVolume& v = c.initialVolumes[i];
    aq_flip_volume& d = a.volumes[i];
    d.shape = v.shapeType == ShapeType::Sphere ? AQ_SHAPE_SPHERE : AQ_SHAPE_BOX;
    d.box_min[0] = v.boxMin.x, d.box_min[1] = v.boxMin.y, d.box_min[2] = v.boxMin.z;
    d.box_max[0] = v.boxMax.x, d.box_max[1] = v.boxMax.y, d.box_max[2] = v.boxMax.z;
    d.center[0] = v.ballCenter.x, d.center[1] = v.ballCenter.y, d.center[2] = v.ballCenter.z;
    d.radius = v.ballRadius;
    d.phase = v.phase;
  }
  return a;
}

class Fluid : public Fluid_3D {
 public:
  Fluid() {}
  ~Fluid() override {
    if (handle) aq_flip_destroy(handle);
  }
  Fluid(const Fluid&) = delete;
  Fluid& operator=(const Fluid&) = delete;

  // Fluid::init (Fluid/Fluid_3D_FLIP.cu:435-505).
  void init(FluidConfig cfg) override {
    if (cfg.phaseCount > 4) throw "Only supports up to 4 phases";  // Fluid_3D_FLIP.cu:449-451
    config = cfg;
    if (handle) {
      aq_flip_destroy(handle);
      handle = nullptr;
    }
    aq_flip_config a = toAbi(cfg);
    check(aq_flip_create(&a, &handle));
    timestep = cfg.FLIP.timestep;
    sizeX = cfg.FLIP.sizeX;
    sizeY = cfg.FLIP.sizeY;
    sizeZ = cfg.FLIP.sizeZ;
    gridPhysicalSize = cfg.gridPhysicalSize;
    cellCount = (sizeX + 1) * (sizeY + 1) * (sizeZ + 1);
    cellPhysicalSize = gridPhysicalSize / (float)sizeY;
    particleCount = (int)aq_flip_particle_count(handle);
    physicalTime = 0;
  }

  // Fluid::simulationStep (Fluid/Fluid_3D_FLIP.cu:316-352).
  void simulationStep() override {
    check(aq_flip_step(handle));
    physicalTime = (float)aq_flip_physical_time(handle);
  }

  void draw(const DrawCommand&) override {}

  float3 getCenter() override {
    float c[3];
    check(aq_flip_get_center(handle, c));
    return make_float3(c[0], c[1], c[2]);
  }
  float getContainerSize() override { return aq_flip_container_size(handle); }

  aq_flip* abiHandle() const { return handle; }

  float timestep = 0;
  int sizeX = 0, sizeY = 0, sizeZ = 0;
  float gridPhysicalSize = 10.f;
  int cellCount = 0;
  float cellPhysicalSize = 0;
  int particleCount = 0;
  const int particlesPerCell = 8;
  FluidConfig config;

 private:
  aq_flip* handle = nullptr;
};

}  // namespace Fluid_3D_FLIP
