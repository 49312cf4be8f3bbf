// aqflip_driver.cpp -- headless replacement of the reference's app loop (main.cu:34-213)
// for the FLIP solver: builds a FluidConfig, calls init(), then simulationStep() in
// a loop, through the drop-in C++ surface in Fluid_3D_FLIP.hpp. Prints one JSON line.
//
//   aqflip_driver [--n N] [--steps K] [--pressure-iters P] [--dt DT] [--multiphase]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "Fluid_3D_FLIP.hpp"

int main(int argc, char** argv) {
  int n = 16, steps = 10, piters = 60;
  float dt = 0.01f;
  bool multiphase = false;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--n") && i + 1 < argc) n = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--steps") && i + 1 < argc) steps = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--pressure-iters") && i + 1 < argc) piters = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--dt") && i + 1 < argc) dt = (float)std::atof(argv[++i]);
    else if (!std::strcmp(argv[i], "--multiphase")) multiphase = true;
  }
  FluidConfig config;
  config.FLIP.sizeX = config.FLIP.sizeY = config.FLIP.sizeZ = n;
  config.FLIP.timestep = dt;
  config.FLIP.pressureIterations = piters;
  config.FLIP.particleJitter = 1.0f;
  config.phaseCount = multiphase ? 2 : 1;
  config.pressureInit = AQ_PRESSURE_ZERO;
  config.seed = 0xA9A00000ull + 1;
  config.initialVolumes[0].boxMax = make_float3(1, 0.25f, 1);
  config.initialVolumes[1].ballCenter = make_float3(0.5f, 0.7f, 0.5f);
  config.initialVolumes[1].ballRadius = multiphase ? 0.15f : 0.2f;
  config.initialVolumes[1].phase = multiphase ? 1 : 0;

  std::shared_ptr<Fluid_3D> fluid;
  try {
    fluid = std::make_shared<Fluid_3D_FLIP::Fluid>();
    fluid->init(config);
    auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < steps; ++s) fluid->simulationStep();
    auto t1 = std::chrono::steady_clock::now();
    auto* flip = static_cast<Fluid_3D_FLIP::Fluid*>(fluid.get());
    double rel = 0;
    Fluid_3D_FLIP::check(aq_flip_residual_fp64(flip->abiHandle(), &rel));
    aq_flip_stats st;
    Fluid_3D_FLIP::check(aq_flip_volume_stats(flip->abiHandle(), &st));
    const double sec = std::chrono::duration<double>(t1 - t0).count();
    float3 c = fluid->getCenter();
    std::printf(
        "{\"n\": %d, \"steps\": %d, \"particles\": %d, \"physical_time\": %.6f, "
        "\"steps_per_sec\": %.3f, \"residual\": %.6e, \"fluid_cells\": %lld, "
        "\"center\": [%g, %g, %g], \"container\": %g}\n",
        n, steps, flip->particleCount, fluid->physicalTime, steps / sec, rel,
        (long long)st.fluid_cells, c.x, c.y, c.z, fluid->getContainerSize());
  } catch (const char* msg) {
    std::fprintf(stderr, "error: %s\n", msg);
    return 2;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
