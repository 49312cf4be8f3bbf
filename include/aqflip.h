/*
 * aqflip.h -- C ABI of the MI355X-native FLIP simulation step.
 *
 * The library (libaqflip.so) owns all device memory behind an opaque handle and
 * runs every kernel of the step on one HIP stream per handle. Callers own every
 * host buffer they pass in. All entry points are synchronous at return, return an
 * aq_status (never exit the process) and leave a message for aq_last_error().
 * One handle per host thread; handles share no state.
 *
 * Reference surface this replaces (paths relative to the AmesingFlank/Aquarius tree):
 *   class Fluid / Fluid_3D            Fluid/Fluid.h:4-12, Fluid/Fluid_3D.cuh:7-13
 *   Fluid_3D_FLIP::Fluid::init        Fluid/Fluid_3D_FLIP.cu:435-505
 *   Fluid_3D_FLIP::Fluid::simulationStep  Fluid/Fluid_3D_FLIP.cu:316-352
 *   MAC_Grid_3D fields and extents    Fluid/MAC_Grid_3D.cuh:23-63, Fluid/MAC_Grid_3D.cu:43-75
 *   FluidConfig / ConfigFLIP          Fluid/FluidConfig.cuh:16-97
 */
#ifndef AQFLIP_H
#define AQFLIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. Replace the reference's exit(EXIT_FAILURE) on any CUDA error
 * (Common/GpuCommons.h:118-142) and its `throw "Only supports up to 4 phases"`
 * (Fluid/Fluid_3D_FLIP.cu:449-451). */
typedef enum aq_status {
  AQ_OK = 0,
  AQ_EINVAL = 1, /* bad argument or configuration */
  AQ_ENOMEM = 2, /* device allocation failed */
  AQ_EHIP = 3,   /* HIP runtime error (launch, copy, device) */
  AQ_ESTALL = 4  /* pressure solve hit its iteration cap or stalled */
} aq_status;

/* Initial fluid volume: InitializationVolume (Fluid/FluidConfig.cuh:20-39).
 * Coordinates are normalised to the unit cube; volumes outside [0,1] are skipped
 * silently, as Fluid/Fluid_3D_FLIP.cu:542-556,571-574 does. */
enum { AQ_SHAPE_SPHERE = 0, AQ_SHAPE_BOX = 1 };
typedef struct aq_flip_volume {
  int32_t shape;     /* AQ_SHAPE_SPHERE or AQ_SHAPE_BOX (ShapeType, FluidConfig.cuh:16-18) */
  float box_min[3];  /* boxMin */
  float box_max[3];  /* boxMax */
  float center[3];   /* ballCenter */
  float radius;      /* ballRadius, in units of the domain's y extent */
  int32_t phase;     /* 0..3 selects the one-hot volume fraction (Fluid_3D_FLIP.cuh:31-48) */
} aq_flip_volume;

#define AQ_MAX_VOLUMES 16

/* Pressure initial guess for each step's Jacobi solve. The reference never
 * clears `pressure` between steps (Fluid/Fluid_3D_FLIP.cu:287-302), so it
 * warm-starts from the previous step: AQ_PRESSURE_PREVIOUS. */
enum { AQ_PRESSURE_ZERO = 0, AQ_PRESSURE_PREVIOUS = 1 };

/* POD mirror of FluidConfig + ConfigFLIP (Fluid/FluidConfig.cuh:41-97). */
typedef struct aq_flip_config {
  int32_t size[3];            /* sizeX, sizeY, sizeZ (cells) */
  float dt;                   /* timestep */
  float flip_coeff;           /* FLIPcoeff (PIC/FLIP blend) */
  float particle_jitter;      /* particleJitter */
  float grid_physical_size;   /* Fluid::gridPhysicalSize = 10 (Fluid_3D_FLIP.cuh:62) */
  int32_t pressure_iters;     /* pressureIterations */
  int32_t diffusion_iters;    /* diffusionIterations */
  int32_t phase_count;        /* phaseCount, 1..4 */
  float gravity[3];           /* gravity */
  float phase_densities[4];   /* phaseDensities */
  float diffusion_coeff;      /* diffusionCoeff */
  int32_t pressure_init;      /* AQ_PRESSURE_ZERO or AQ_PRESSURE_PREVIOUS */
  int32_t ref_boundary_coverage; /* 1: fixBoundaryX/Y/Z cover only the pairs the
                                    reference's launch reaches (Fluid_3D_common.cu:13-41);
                                    0: every wall pair */
  uint64_t seed;              /* particle jitter generator seed (replaces rand()) */
  int32_t n_volumes;
  aq_flip_volume volumes[AQ_MAX_VOLUMES];
} aq_flip_config;

/* Grid fields, exported in the reference extents (Fluid/MAC_Grid_3D.cu:53-71).
 * Host layout: x-major, z fastest, i.e. element (x,y,z) of an extent (EX,EY,EZ)
 * is at ((x*EY)+y)*EZ+z; 4-wide fields are interleaved (.x,.y,.z,.w).
 *   id                         type   extent     reference field
 */
enum {
  AQ_FIELD_CONTENT = 0,        /* int32  (N+1)^3  content (AIR 0 / FLUID 1 / SOLID 2) */
  AQ_FIELD_VELOCITY = 1,       /* float4 (N+1)^3  velocity (P2G result) */
  AQ_FIELD_NEW_VELOCITY = 2,   /* float4 (N+1)^3  newVelocity */
  AQ_FIELD_PRESSURE = 3,       /* float  N^3      pressure */
  AQ_FIELD_DIVERGENCE = 4,     /* float  N^3      divergence */
  AQ_FIELD_DENSITY = 5,        /* float  N^3      density */
  AQ_FIELD_PARTICLE_COUNT = 6, /* int32  N^3      particleCount (= cellEnd - cellStart) */
  AQ_FIELD_HAS_VELOCITY = 7,   /* int4   N^3      hasVelocity */
  AQ_FIELD_VOLUME_FRACTIONS = 8,     /* float4 N^3 volumeFractions */
  AQ_FIELD_NEW_VOLUME_FRACTIONS = 9, /* float4 N^3 newVolumeFractions */
  AQ_FIELD_COUNT = 10
};

/* Stages of simulationStep (Fluid/Fluid_3D_FLIP.cu:316-352), runnable one at a
 * time for per-kernel parity checks. */
enum {
  AQ_STAGE_G2P = 0,              /* transferToParticles      Fluid_3D_FLIP.cu:189-227,380-384 */
  AQ_STAGE_ADVECT = 1,           /* moveParticles (RK3)       Fluid_3D_FLIP.cu:229-285,386-394 */
  AQ_STAGE_TRANSFER_TO_GRID = 2, /* hash, sort, P2G          Fluid_3D_FLIP.cu:363-378 */
  AQ_STAGE_GRAVITY = 3,          /* applyGravity             Fluid_3D_kernels.cu:4-39 */
  AQ_STAGE_BOUNDARY = 4,         /* fixBoundary              Fluid_3D_common.cu:13-41 */
  AQ_STAGE_DIVERGENCE = 5,       /* countParticles + computeDivergence  Fluid_3D_FLIP.cu:168-185, Fluid_3D_kernels.cu:426-458 */
  AQ_STAGE_PRESSURE = 6,         /* solvePressureJacobi      Fluid_3D_common.cu:48-55 */
  AQ_STAGE_PROJECT = 7,          /* updateVelocityWithPressure  Fluid_3D_kernels.cu:118-207 */
  AQ_STAGE_EXTRAPOLATE = 8,      /* extrapolateVelocity (3 passes)  Fluid_3D_common.cu:65-82 */
  AQ_STAGE_DIFFUSION = 9,        /* solveDiffusionJacobi     Fluid_3D_common.cu:85-91 */
  /* Pressure and diffusion sweeps paired in one launch inside a fixed-sweep step (the
   * two chains are independent); a timing bucket only, not runnable on its own. */
  AQ_STAGE_PRESSURE_DIFFUSION = 10,
  /* The P2G kernel of transferToGrid, timed apart from the hashing and sorting
   * inside a step; a timing bucket only (AQ_STAGE_TRANSFER_TO_GRID runs both). */
  AQ_STAGE_P2G = 11,
  AQ_STAGE_COUNT = 12
};

typedef struct aq_flip aq_flip; /* opaque handle */

/* Reference defaults of FluidConfig (Fluid/FluidConfig.cuh:20-97) plus the
 * choices this library adds: pressure_init PREVIOUS, reference boundary
 * coverage, seed 0. */
aq_status aq_flip_config_default(aq_flip_config* cfg);

/* Host seeding: createSquareFluid / createSphereFluid / createParticlesAt
 * (Fluid/Fluid_3D_FLIP.cu:508-594) with a seeded SplitMix64 stream in place of
 * rand() (Common/Commons.h:3-5). pos is n*3 floats, frac4 n*4 floats. Pass
 * pos == NULL to query the count in *n_out. */
aq_status aq_flip_seed_particles(const aq_flip_config* cfg, float* pos, float* frac4,
                                 int64_t capacity, int64_t* n_out);

/* = Fluid::init (Fluid/Fluid_3D_FLIP.cu:435-505) without the GL renderers and
 * the mesher: allocate, seed, upload, fixBoundary, transferToGrid. */
aq_status aq_flip_create(const aq_flip_config* cfg, aq_flip** out);

/* Like aq_flip_create but with caller-supplied particles (n*3 pos, n*3 vel,
 * n*4 frac4) in place of seeding; then fixBoundary and transferToGrid as init.
 * Positions must lie in [0, size*dx) on every axis. */
aq_status aq_flip_create_with_particles(const aq_flip_config* cfg, const float* pos,
                                        const float* vel, const float* frac4, int64_t n,
                                        aq_flip** out);

void aq_flip_destroy(aq_flip* h);

/* = Fluid::simulationStep (Fluid/Fluid_3D_FLIP.cu:316-352). */
aq_status aq_flip_step(aq_flip* h);

/* Run `count` steps back to back without host synchronisation in between. */
aq_status aq_flip_steps(aq_flip* h, int32_t count);

/* Run one AQ_STAGE_* on the current state (per-kernel parity). */
aq_status aq_flip_run_stage(aq_flip* h, int32_t stage);

/* Particles, in the library's current (sorted) order. Any pointer may be NULL. */
int64_t aq_flip_particle_count(const aq_flip* h);
aq_status aq_flip_get_particles(aq_flip* h, float* pos, float* vel, float* frac4);
/* Replace particle state (same count as the handle holds). No grid work is done. */
aq_status aq_flip_set_particles(aq_flip* h, const float* pos, const float* vel,
                                const float* frac4, int64_t n);

/* Grid fields in the reference extents and layout (see AQ_FIELD_*). `bytes`
 * must equal the field's exported size, which aq_flip_field_bytes returns. */
size_t aq_flip_field_bytes(const aq_flip* h, int32_t field);
aq_status aq_flip_get_field(aq_flip* h, int32_t field, void* dst, size_t bytes);
aq_status aq_flip_set_field(aq_flip* h, int32_t field, const void* src, size_t bytes);

/* Spatial-hash state after the last sort (performSpatialHashing2,
 * Fluid/Fluid_3D_common.cuh:27-56): cell_start/cell_end in the reference
 * convention (inclusive end, -1 when empty, (X+1)(Y+1)(Z+1) entries), per-particle
 * hash in sorted order and the permutation (sorted slot -> previous index).
 * Any pointer may be NULL. */
aq_status aq_flip_get_cells(aq_flip* h, int32_t* cell_start, int32_t* cell_end,
                            int32_t* hash, int32_t* perm);

/* FP64 relative residual ||b - A p||_2 / ||b||_2 over FLUID cells of the
 * current pressure, b = -divergence*density, A the operator whose fixed point
 * the Jacobi iteration of Fluid/Fluid_3D_kernels.cu:460-505 is. Computed on the
 * GPU. */
aq_status aq_flip_residual_fp64(aq_flip* h, double* rel_l2);

/* FP64 volume and mass checks, computed on the GPU. */
typedef struct aq_flip_stats {
  int64_t particles;      /* particle count */
  int64_t fluid_cells;    /* interior cells marked FLUID */
  double phase_sum[4];    /* sum over particles of each volume fraction */
  double mean_ppc;        /* particles / fluid_cells */
  double pos_min[3], pos_max[3]; /* particle position bounds */
} aq_flip_stats;
aq_status aq_flip_volume_stats(aq_flip* h, aq_flip_stats* out);

/* Pressure solve to a tolerance on the current divergence: Jacobi sweeps from
 * the current pressure until the FP64 relative residual is <= tol, checked
 * every check_every sweeps, at most max_iters sweeps. Returns AQ_ESTALL when
 * the cap is hit. No counterpart in the reference (which runs a fixed count). */
aq_status aq_flip_solve_pressure(aq_flip* h, double tol, int32_t max_iters,
                                 int32_t check_every, int32_t* iters_out,
                                 double* rel_out);

/* The same system solved by multigrid-preconditioned conjugate gradients in FP64
 * (one geometric V-cycle per iteration), the accuracy reference and second
 * fallback beside Jacobi; the reference's counterpart is the legacy sparse path,
 * CSR assembly (Fluid/Legacy/PressureEquation.cuh:26-211,239-328) solved by
 * AMGX-preconditioned CG (Fluid/Legacy/SPD_Solver.h:102-160). Starts from the
 * current pressure when from_current != 0, else from zero; iterates until the CG
 * residual is <= tol (checked every check_every iterations) or max_iters; stores
 * the result as FP32. rel_out is the FP64 relative residual of that stored FP32
 * field. Returns AQ_ESTALL when the CG residual did not reach tol. */
aq_status aq_flip_solve_pressure_mgpcg(aq_flip* h, double tol, int32_t max_iters,
                                       int32_t check_every, int32_t from_current,
                                       int32_t* iters_out, double* rel_out);

/* Per-stage device timing with HIP events on the handle's stream. When enabled,
 * every stage of aq_flip_step records events; aq_flip_get_timing returns, per
 * AQ_STAGE_* (arrays of AQ_STAGE_COUNT) since the last reset, the accumulated
 * milliseconds and the number of passes: sweeps for the pressure, diffusion and
 * paired stages, 3 per extrapolation, 1 per run of any other stage. */
aq_status aq_flip_set_timing(aq_flip* h, int32_t enable);
aq_status aq_flip_get_timing(aq_flip* h, double* stage_ms, int64_t* stage_calls);
aq_status aq_flip_reset_timing(aq_flip* h);

/* Paired pressure + diffusion sweeps (AQ_STAGE_PRESSURE_DIFFUSION) whose diffusion
 * part was computed, since the handle was created (32-bit device counter). A paired
 * sweep skips its diffusion part when the previous diffusion sweep of the step changed
 * no bit: every later sweep would recompute the same values from the same inputs.
 * No counterpart in the reference, which always computes every sweep. */
aq_status aq_flip_paired_diffusion_computed(aq_flip* h, int64_t* count);

/* Simulated time (Fluid::physicalTime, Fluid/Fluid.h:11). */
double aq_flip_physical_time(const aq_flip* h);
/* Domain centre and size: Fluid_3D::getCenter / getContainerSize
 * (Fluid/Fluid_3D_FLIP.cu:728-733). center is 3 floats. */
aq_status aq_flip_get_center(const aq_flip* h, float* center);
float aq_flip_container_size(const aq_flip* h);

/* Checkpoint and resume (no counterpart in the reference, which has none). The
 * file is a versioned little-endian snapshot: magic "AQFLIPSN", format version,
 * the configuration field by field, the step state (particle count, phase channels,
 * physical time, ping-pong selectors) and tagged sections holding the particles in
 * sorted order, every grid field in the library's padded layout and the cell bins
 * -- all that a later aq_flip_step or export reads. aq_flip_load returns a handle
 * whose steps and exports are bit-identical to those of the saved one. The solver
 * mode (aq_flip_set_solver) is not part of the snapshot. */
#define AQ_SNAPSHOT_VERSION 1
aq_status aq_flip_save(aq_flip* h, const char* path);
aq_status aq_flip_load(const char* path, aq_flip** out);

/* Thread-local message of the last failing call on this thread. */
const char* aq_last_error(void);

/* Library version string. */
const char* aq_flip_version(void);

#ifdef __cplusplus
}
#endif

#endif /* AQFLIP_H */
