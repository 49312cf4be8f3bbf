// flip_api.cpp -- C ABI of libaqflip (include/aqflip.h).
//
// Host orchestration of Fluid_3D_FLIP::Fluid (Fluid/Fluid_3D_FLIP.cu:316-505) for
// one GPU: every stage is enqueued on the handle's own HIP stream with no host
// synchronisation inside a step (the reference synchronises about 30 times per
// step); entry points synchronise once before returning. Errors come back as
// aq_status with a thread-local message instead of exit().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/aqflip.h"
#include "../../include/aqflip_unet.h"
#include "aqflip_internal.h"
#include "flip_view.h"
#include "mgpcg.h"

namespace aqf {
int64_t seed_scene(const aq_flip_config& cfg, std::vector<float>* pos, std::vector<float>* frac);
}

using namespace aqf;

namespace {

thread_local std::string g_err;

aq_status fail(aq_status s, const std::string& msg) {
  g_err = msg;
  return s;
}

#define HIP_TRY(expr)                                                                         \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail(e_ == hipErrorOutOfMemory ? AQ_ENOMEM : AQ_EHIP,                            \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                         \
  } while (0)

#define AQ_TRY(expr)              \
  do {                            \
    aq_status s_ = (expr);        \
    if (s_ != AQ_OK) return s_;   \
  } while (0)

int boundary_coverage(int A, int B, bool ref) {
  const int pairs = (A + 1) * (B + 1);
  if (!ref) return pairs;
  const int total = A * B;
  const int nt = std::min(1024, total);
  // divUp (Common/Commons.h:7-13)
  const int blocks = nt == 0 ? 1 : (total % nt != 0 ? total / nt + 1 : total / nt);
  return (int)std::min<long long>((long long)blocks * nt, pairs);
}

}  // namespace

struct aq_flip {
  aq_flip_config cfg{};
  Geom G{};
  int K = 1;
  long long Np = 0;
  hipStream_t stream = nullptr;
  // Diffusion of the phase fractions depends on neither pressure nor velocities, so
  // inside a step its sweeps run on a second stream, concurrently with the pressure
  // solve, projection and extrapolation (fork after the divergence pass, join before
  // the step ends).
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool overlap = true;
  // With the fixed-sweep solver, diffusion sweep i runs in the same launch as
  // pressure sweep i (k_jacobi_diffusion) instead of on the second stream.
  bool fuse = true;
  // Paired diffusion sweeps after one that changed no bit skip their identical work.
  bool skip_fixed = true;
  int listed_grid = 0;  // 0: the default listed sweep grids; n: at most n workgroups
  std::vector<void*> allocs;

  Particles P[2]{};
  int cur = 0;
  Grid gr{};
  int nv = 0, hv = 0, pb = 0, nf = 0;  // current buffer of each ping-pong pair
  // air_clean[i]: pressure buffer i holds 0 on every AIR cell of the current content.
  bool air_clean[2] = {false, false};
  Bins B{};
  StepParams sp{};
  int nparts_red = 0;
  double* red_partial = nullptr;
  double* red_out = nullptr;
  float physicalTime = 0.f;

  // Pressure solve mode (fixed sweeps as the reference, or to a tolerance).
  aq_flip_solver solver{AQ_SOLVE_FIXED, 1e-4, 100000, 50, 20, 0};
  aq_unet* net = nullptr;
  aq_flip_solver_stats sstats{};
  float* p_keep = nullptr;  // the iterate before a U-Net correction (allocated on first use)

  bool timing = false;
  double stage_ms[AQ_STAGE_COUNT]{};
  int64_t stage_calls[AQ_STAGE_COUNT]{};
  // Stage timing: event pairs recorded on this handle's stream while a step is
  // enqueued, read back only at the synchronisation point that ends the call.
  struct Mark {
    int stage;
    hipEvent_t a, b;
    int64_t passes;
  };
  // Captured steps: one executable graph per ping-pong state at step entry (odd
  // pass counts alternate buffers between steps), with the state it leaves.
  struct StepGraph {
    int key[7];
    int next[7];
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  std::vector<StepGraph> graphs;
  bool use_graphs = true;
  bool nvo_valid = false;  // gr.nvo matches newVelocity and velocity (see graph_step)
  // Tolerance solves: k Jacobi sweeps plus the FP64 residual kernels between two
  // host checks, captured once per (pressure buffer, AIR-clean state, k).
  struct SweepGraph {
    int key[4];   // pb, air_clean[0], air_clean[1], k
    int next[3];  // pb, air_clean[0], air_clean[1] afterwards
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  std::vector<SweepGraph> sweep_graphs;
  static constexpr size_t kMaxSweepGraphs = 16;
  // Multigrid-preconditioned CG (allocated on first use).
  struct MgState {
    MgHierarchy H;
    double *b = nullptr, *x = nullptr, *r = nullptr, *z = nullptr, *d = nullptr, *q = nullptr;
    double *part = nullptr, *S = nullptr;
    int nparts = 256;
  };
  MgState* mg = nullptr;
  std::vector<hipEvent_t> event_pool;
  std::vector<Mark> pending;
  size_t pool_used = 0;
  hipEvent_t open_event = nullptr;

  ~aq_flip() {
    if (stream) hipStreamSynchronize(stream);
    delete mg;
    for (StepGraph& g : graphs) {
      hipGraphExecDestroy(g.exec);
      hipGraphDestroy(g.graph);
    }
    for (SweepGraph& g : sweep_graphs) {
      hipGraphExecDestroy(g.exec);
      hipGraphDestroy(g.graph);
    }
    for (void* p : allocs) hipFree(p);
    for (hipEvent_t e : event_pool) hipEventDestroy(e);
    if (side) hipStreamSynchronize(side);
    if (ev_fork) hipEventDestroy(ev_fork);
    if (ev_join) hipEventDestroy(ev_join);
    if (side) hipStreamDestroy(side);
    if (stream) hipStreamDestroy(stream);
  }

  aq_status alloc(void** p, size_t bytes) {
    HIP_TRY(hipMalloc(p, bytes));
    allocs.push_back(*p);
    HIP_TRY(hipMemsetAsync(*p, 0, bytes, stream));
    return AQ_OK;
  }
  template <typename T>
  aq_status allocT(T** p, size_t n) {
    return alloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
  }

  hipEvent_t next_event() {
    if (pool_used == event_pool.size()) {
      hipEvent_t e = nullptr;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      event_pool.push_back(e);
    }
    return event_pool[pool_used++];
  }
  void begin(int st) {
    (void)st;
    if (!timing) return;
    open_event = next_event();
    if (open_event) hipEventRecord(open_event, stream);
  }
  void end(int st, int64_t passes = 1) {
    if (!timing || !open_event) return;
    hipEvent_t e = next_event();
    if (!e) return;
    hipEventRecord(e, stream);
    pending.push_back({st, open_event, e, passes});
    open_event = nullptr;
  }
  // Called after the stream has been synchronised.
  aq_status collect() {
    for (const Mark& m : pending) {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, m.a, m.b));
      stage_ms[m.stage] += ms;
      stage_calls[m.stage] += m.passes;
    }
    pending.clear();
    pool_used = 0;
    return AQ_OK;
  }
};

namespace {

aq_status validate(const aq_flip_config* c) {
  if (!c) return fail(AQ_EINVAL, "null config");
  for (int a = 0; a < 3; ++a)
    if (c->size[a] < 1 || c->size[a] > 1024)
      return fail(AQ_EINVAL, "grid size must be in [1, 1024] on every axis");
  if ((long long)c->size[0] * c->size[1] * c->size[2] > (1ll << 30))
    return fail(AQ_EINVAL, "grid too large");
  if (c->phase_count > 4) return fail(AQ_EINVAL, "Only supports up to 4 phases");
  if (c->phase_count < 1) return fail(AQ_EINVAL, "phase_count must be >= 1");
  if (c->n_volumes < 0 || c->n_volumes > AQ_MAX_VOLUMES)
    return fail(AQ_EINVAL, "n_volumes out of range");
  if (!(c->grid_physical_size > 0) || !std::isfinite(c->grid_physical_size))
    return fail(AQ_EINVAL, "grid_physical_size must be positive");
  if (!std::isfinite(c->dt) || !std::isfinite(c->flip_coeff))
    return fail(AQ_EINVAL, "dt and flip_coeff must be finite");
  if (c->pressure_iters < 0 || c->diffusion_iters < 0)
    return fail(AQ_EINVAL, "iteration counts must be >= 0");
  if (c->pressure_init != AQ_PRESSURE_ZERO && c->pressure_init != AQ_PRESSURE_PREVIOUS)
    return fail(AQ_EINVAL, "pressure_init must be AQ_PRESSURE_ZERO or AQ_PRESSURE_PREVIOUS");
  return AQ_OK;
}

// Number of volume-fraction channels in use: phases 0..K-1 can be non-zero; the
// remaining channels are provably zero throughout the step.
int channels_for(const aq_flip_config& c, const float* frac4, long long n) {
  int K = std::max(1, std::min(4, c.phase_count));
  for (int i = 0; i < c.n_volumes; ++i)
    if (c.volumes[i].phase >= 0 && c.volumes[i].phase < 4) K = std::max(K, c.volumes[i].phase + 1);
  if (frac4)
    for (long long i = 0; i < n; ++i)
      for (int a = 3; a >= K; --a)
        if (frac4[4 * i + a] != 0.f) {
          K = a + 1;
          break;
        }
  return K;
}

// Positions must hash into the grid (calcHashImpl would index out of range otherwise).
aq_status check_positions(const aq_flip* h, const float* pos, long long n) {
  const int S[3] = {h->G.X, h->G.Y, h->G.Z};
  for (long long i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      const float v = pos[3 * i + a];
      if (!(v >= 0.f) || (int)(v / h->G.h) >= S[a])
        return fail(AQ_EINVAL, "particle " + std::to_string(i) + " lies outside the grid");
    }
  return AQ_OK;
}

aq_status setup(aq_flip* h, const aq_flip_config* cfg, long long Np) {
  h->cfg = *cfg;
  Geom& G = h->G;
  G.X = cfg->size[0];
  G.Y = cfg->size[1];
  G.Z = cfg->size[2];
  G.EY = G.Y + 1;
  G.EZ = ((G.Z + 1) + 7) / 8 * 8;
  G.sx = (long long)G.EY * G.EZ;
  G.ncell = (long long)(G.X + 1) * G.sx;
  G.ncells_hash = G.X * G.Y * G.Z;
  G.nfaces = (G.X + 1) * (G.Y + 1) * (G.Z + 1);
  G.h = cfg->grid_physical_size / (float)G.Y;  // Fluid_3D_FLIP.cu:444
  G.inv_h = 1.0 / (double)G.h;
  G.divZ = make_fastdiv((uint32_t)G.Z);
  G.divYZ = make_fastdiv((uint32_t)(G.Y * G.Z));
  G.divZ1 = make_fastdiv((uint32_t)(G.Z + 1));
  G.divYZ1 = make_fastdiv((uint32_t)((G.Y + 1) * (G.Z + 1)));
  h->Np = Np;

  StepParams& sp = h->sp;
  sp.dt = cfg->dt;
  sp.flip = cfg->flip_coeff;
  for (int a = 0; a < 3; ++a) sp.g[a] = cfg->gravity[a];
  for (int a = 0; a < 4; ++a) sp.rho[a] = cfg->phase_densities[a];
  sp.lambda = (cfg->diffusion_coeff * cfg->dt) / (G.h * G.h);  // Fluid_3D_common.cu:86
  sp.inv_cc = 1.0 / (double)(1.f + 2.f * sp.lambda);             // Fluid_3D_kernels.cu:538
  const bool ref = cfg->ref_boundary_coverage != 0;
  sp.coverX = boundary_coverage(G.Y, G.Z, ref);
  sp.coverY = boundary_coverage(G.X, G.Z, ref);
  sp.coverZ = boundary_coverage(G.X, G.Y, ref);

  HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
  for (int b = 0; b < 2; ++b) {
    AQ_TRY(h->allocT(&h->P[b].a, (size_t)Np));
    AQ_TRY(h->allocT(&h->P[b].b, (size_t)Np));
    AQ_TRY(h->allocT(&h->P[b].c, (size_t)Np));
  }
  const size_t nc = (size_t)G.ncell;
  Grid& gr = h->gr;
  AQ_TRY(h->allocT(&gr.content, nc));
  for (int a = 0; a < 3; ++a) {
    AQ_TRY(h->allocT(&gr.vel[a], nc));
    AQ_TRY(h->allocT(&gr.nvel[0][a], nc));
    AQ_TRY(h->allocT(&gr.nvel[1][a], nc));
  }
  for (int a = 0; a < 3; ++a) AQ_TRY(h->allocT(&gr.nvo[a], nc));
  AQ_TRY(h->allocT(&gr.hasvel[0], nc));
  AQ_TRY(h->allocT(&gr.hasvel[1], nc));
  AQ_TRY(h->allocT(&gr.p[0], nc));
  AQ_TRY(h->allocT(&gr.p[1], nc));
  AQ_TRY(h->allocT(&gr.div, nc));
  AQ_TRY(h->allocT(&gr.rhs, nc));
  AQ_TRY(h->allocT(&gr.density, nc));
  // One flag per 256-cell block plus one slot that stays 0, so the sweep kernels'
  // pairs of blocks never read past the array.
  AQ_TRY(h->allocT(&gr.active, (size_t)(G.ncells_hash + 255) / 256 + 1));
  AQ_TRY(h->allocT(&gr.alist, (size_t)sweep_chunks(G) + 1));
  AQ_TRY(h->allocT(&gr.btab, (size_t)sweep_table_entries(G)));
  set_listed_grids(G, gr, h->listed_grid);
  AQ_TRY(h->allocT(&gr.dchange, 4));
  AQ_TRY(h->allocT(&gr.stepstamp, 1));
  AQ_TRY(h->allocT(&gr.rowstamp, (size_t)(G.X + 1) * (G.Y + 1)));
  for (int a = 0; a < 4; ++a) {
    AQ_TRY(h->allocT(&gr.vf[a], nc));
    AQ_TRY(h->allocT(&gr.nvf[0][a], nc));
    AQ_TRY(h->allocT(&gr.nvf[1][a], nc));
  }
  const int nch = G.ncells_hash;
  AQ_TRY(h->allocT(&h->B.count, (size_t)nch));
  AQ_TRY(h->allocT(&h->B.offsets, (size_t)nch + 1));
  AQ_TRY(h->allocT(&h->B.partials, (size_t)(nch + 4095) / 4096 + 1));
  AQ_TRY(h->allocT(&h->B.hash, (size_t)Np));
  AQ_TRY(h->allocT(&h->B.rank, (size_t)Np));
  AQ_TRY(h->allocT(&h->B.perm, (size_t)Np));
  AQ_TRY(h->allocT(&h->B.multi, (size_t)nch));  // zeroed here; k_segsort clears what it reads
  h->nparts_red = 512;  // volume statistics: 512 partial rows of 16 columns
  // (the residual uses two columns per sweep chunk)
  AQ_TRY(h->allocT(&h->red_partial, std::max<size_t>((size_t)h->nparts_red * 16, 2 * (size_t)sweep_chunks(G))));
  AQ_TRY(h->allocT(&h->red_out, 16));
  return AQ_OK;
}

void refresh_rhs(aq_flip* h) { launch_refresh_rhs(h->G, h->gr, h->stream); }

aq_status upload_particles(aq_flip* h, const float* pos, const float* vel, const float* frac4) {
  const long long n = h->Np;
  std::vector<float4> a((size_t)n), b((size_t)n), c((size_t)n);
  for (long long i = 0; i < n; ++i) {
    const float* p = pos + 3 * i;
    float v[3] = {0.f, 0.f, 0.f}, f[4] = {0.f, 0.f, 0.f, 0.f};
    if (vel)
      for (int k = 0; k < 3; ++k) v[k] = vel[3 * i + k];
    if (frac4)
      for (int k = 0; k < 4; ++k) f[k] = frac4[4 * i + k];
    a[i] = make_float4(p[0], p[1], p[2], v[0]);
    b[i] = make_float4(v[1], v[2], f[0], f[1]);
    c[i] = make_float4(f[2], f[3], 0.f, 0.f);
  }
  const Particles& P = h->P[h->cur];
  HIP_TRY(hipMemcpyAsync(P.a, a.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipMemcpyAsync(P.b, b.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipMemcpyAsync(P.c, c.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return AQ_OK;
}

// ---- stages --------------------------------------------------------------

// New FLUID/AIR marking: AIR cells of either pressure buffer may hold old values.
void content_changed(aq_flip* h) { h->air_clean[0] = h->air_clean[1] = false; }

// hash is already binned (count filled): scan, scatter, segment sort, gather, P2G.
// Inside a step the P2G kernel is timed in its own bucket (split = true).
void transfer_after_count(aq_flip* h, bool split = false) {
  launch_scan(h->G, h->B, h->stream);
  launch_scatter_sort_gather(h->G, h->P[h->cur], h->P[h->cur ^ 1], h->Np, h->K, h->B, h->stream);
  h->cur ^= 1;
  if (split) {
    h->end(AQ_STAGE_TRANSFER_TO_GRID);
    h->begin(AQ_STAGE_P2G);
  }
  launch_p2g(h->G, h->P[h->cur], h->Np, h->K, h->B, h->gr, h->nv, h->hv, h->nf, h->sp, h->stream);
  if (split) h->end(AQ_STAGE_P2G);
  content_changed(h);
}

void clear_bins(aq_flip* h) { launch_fill_u32(h->B.count, 0u, h->G.ncells_hash, h->stream); }

void stage_transfer(aq_flip* h) {
  clear_bins(h);
  launch_hash_count(h->G, h->P[h->cur], h->Np, h->B, h->stream);
  transfer_after_count(h);
}

// n Jacobi sweeps. Every sweep leaves 0 on all AIR cells of its output buffer, so
// a sweep whose output buffer already holds those zeros may skip AIR cells and
// blocks without FLUID (the result is identical); otherwise it writes every cell.
// from_zero: the sweeps start from the zero guess; the first one does not read the
// input buffer (its neighbours are 0) and writes every cell, so neither buffer needs
// clearing first.
void jacobi_sweeps(aq_flip* h, int n, bool from_zero = false) {
  for (int i = 0; i < n; ++i) {
    const int out = h->pb ^ 1;
    const bool z = from_zero && i == 0;
    launch_jacobi(h->G, h->gr, h->pb, (z || !h->air_clean[out]) ? 1 : 0, z ? 1 : 0, h->stream);
    h->air_clean[out] = true;
    h->pb = out;
  }
}


// Zero pressure start: both buffers cleared.
void zero_pressure(aq_flip* h) {
  launch_fill_pressure_zero(h->G, h->gr, h->pb, h->stream);
  launch_fill_pressure_zero(h->G, h->gr, h->pb ^ 1, h->stream);
  h->air_clean[0] = h->air_clean[1] = true;
}
// The fixed-sweep solve: pressure_iters sweeps from the configured start.
void fixed_sweeps(aq_flip* h) {
  const bool zero = h->cfg.pressure_init == AQ_PRESSURE_ZERO;
  if (zero && h->cfg.pressure_iters == 0) zero_pressure(h);
  jacobi_sweeps(h, h->cfg.pressure_iters, zero);
}

void stage_pressure(aq_flip* h) { fixed_sweeps(h); }

// Inside a step (in_step) the second and third passes visit only the face rows this
// step's P2G stamped (k_extrapolate); the stage on its own runs every pass in full.
// Inside a step the last pass also writes the interleaved (newVelocity, velocity) copy the
// next step's G2P samples (gr.nvo), on the rows it visits.
void stage_extrapolate(aq_flip* h, bool in_step = false) {
  for (int i = 0; i < 3; ++i) {  // maxDistance = 3 (Fluid_3D_common.cu:74)
    launch_extrapolate(h->G, h->gr, h->nv, h->hv, in_step && i > 0, h->stream, in_step && i == 2);
    h->nv ^= 1;
    h->hv ^= 1;
  }
}

// Diffusion sweeps first..diffusion_iters-1 of the step.
void diffusion_sweeps(aq_flip* h, hipStream_t s, int first = 0) {
  for (int i = first; i < h->cfg.diffusion_iters; ++i) {  // the first sweep also copies AIR cells across
    launch_diffusion(h->G, h->gr, h->K, h->nf, h->sp.lambda, h->sp.inv_cc, i == 0 ? 1 : 0, s);
    h->nf ^= 1;
  }
}

// Fixed-sweep pressure solve and the diffusion sweeps together: the first
// min(pressure_iters, diffusion_iters) sweeps of each run pairwise in one launch,
// the rest of the longer chain on its own. Same iterates as the two loops.
void fused_pressure_diffusion(aq_flip* h) {
  const bool zero = h->cfg.pressure_init == AQ_PRESSURE_ZERO;  // first sweep from the zero guess
  const int n = std::min(h->cfg.pressure_iters, h->cfg.diffusion_iters);
  h->begin(AQ_STAGE_PRESSURE_DIFFUSION);
  for (int i = 0; i < n; ++i) {
    const int out = h->pb ^ 1;
    const bool z = zero && i == 0;
    launch_jacobi_diffusion(h->G, h->gr, h->K, h->pb, (z || !h->air_clean[out]) ? 1 : 0, h->nf,
                            h->sp.lambda, h->sp.inv_cc, i == 0 ? 1 : 0, z ? 1 : 0, i,
                            h->skip_fixed ? 1 : 0, h->stream);
    h->air_clean[out] = true;
    h->pb = out;
    h->nf ^= 1;
  }
  h->end(AQ_STAGE_PRESSURE_DIFFUSION, n);
  h->begin(AQ_STAGE_PRESSURE);
  jacobi_sweeps(h, h->cfg.pressure_iters - n);
  h->end(AQ_STAGE_PRESSURE, h->cfg.pressure_iters - n);
  h->begin(AQ_STAGE_DIFFUSION);
  diffusion_sweeps(h, h->stream, n);
  h->end(AQ_STAGE_DIFFUSION, h->cfg.diffusion_iters - n);
}

// inStep: the grid fractions were produced by this step's P2G, so with a single
// channel they are exactly 1 (FLUID) or 0 (AIR) and every sweep is the identity;
// Jacobi with >= 1 sweep has already zeroed AIR pressure.
// Returns the number of diffusion sweeps run.
int stage_diffusion(aq_flip* h, bool inStep) {
  if (h->cfg.diffusion_iters <= 0) return 0;
  if (!inStep || h->cfg.pressure_iters == 0) launch_air_pressure_zero(h->G, h->gr, h->pb, h->stream);
  if (inStep && h->K == 1) return 0;
  diffusion_sweeps(h, h->stream);
  return h->cfg.diffusion_iters;
}

aq_status run_stage(aq_flip* h, int st) {
  if (st == AQ_STAGE_PRESSURE_DIFFUSION || st == AQ_STAGE_P2G)
    return fail(AQ_EINVAL, "timing bucket: runs only inside a step");
  h->nvo_valid = false;
  if (st < 0 || st >= AQ_STAGE_COUNT) return fail(AQ_EINVAL, "unknown stage");
  int64_t passes = 1;
  h->begin(st);
  switch (st) {
    case AQ_STAGE_G2P: launch_g2p(h->G, h->P[h->cur], h->Np, h->K, h->gr, h->nv, h->nf, h->sp, h->stream); break;
    case AQ_STAGE_ADVECT: launch_advect(h->G, h->P[h->cur], h->Np, h->gr, h->nv, h->sp, h->stream); break;
    case AQ_STAGE_TRANSFER_TO_GRID:
      stage_transfer(h);
      refresh_rhs(h);  // content changed: re-mark AIR for Jacobi
      break;
    case AQ_STAGE_GRAVITY: launch_gravity_boundary(h->G, h->gr, h->nv, h->sp, 1, 0, h->stream); break;
    case AQ_STAGE_BOUNDARY: launch_gravity_boundary(h->G, h->gr, h->nv, h->sp, 0, 1, h->stream); break;
    case AQ_STAGE_DIVERGENCE: launch_divergence(h->G, h->gr, h->nv, h->B, h->stream); break;
    case AQ_STAGE_PRESSURE:
      stage_pressure(h);
      passes = h->cfg.pressure_iters;
      break;
    case AQ_STAGE_PROJECT: launch_project(h->G, h->gr, h->nv, h->hv, h->pb, h->stream); break;
    case AQ_STAGE_EXTRAPOLATE:
      stage_extrapolate(h);
      passes = 3;
      break;
    case AQ_STAGE_DIFFUSION: passes = stage_diffusion(h, false); break;
    default: break;
  }
  HIP_TRY(hipGetLastError());
  h->end(st, passes);
  return AQ_OK;
}

// ---- multigrid-preconditioned CG ---------------------------------------------
aq_status mg_setup(aq_flip* h) {
  if (h->mg) return AQ_OK;
  auto* M = new aq_flip::MgState;
  h->mg = M;
  const Geom& G = h->G;
  int X = G.X, Y = G.Y, Z = G.Z;
  for (;;) {  // halve every axis down to a few cells
    MgLevel L{X, Y, Z, (long long)X * Y * Z, nullptr};
    AQ_TRY(h->allocT(&L.fluid, (size_t)L.n));
    float *f, *u, *t;
    AQ_TRY(h->allocT(&f, (size_t)L.n));
    AQ_TRY(h->allocT(&u, (size_t)L.n));
    AQ_TRY(h->allocT(&t, (size_t)L.n));
    M->H.levels.push_back(L);
    M->H.f.push_back(f);
    M->H.u.push_back(u);
    M->H.t.push_back(t);
    if (std::max(X, std::max(Y, Z)) <= 4 || M->H.levels.size() >= 12) break;
    X = (X + 1) / 2, Y = (Y + 1) / 2, Z = (Z + 1) / 2;
  }
  const size_t n0 = (size_t)M->H.levels[0].n;
  for (double** v : {&M->b, &M->x, &M->r, &M->z, &M->d, &M->q}) AQ_TRY(h->allocT(v, n0));
  AQ_TRY(h->allocT(&M->part, (size_t)M->nparts * 2));
  AQ_TRY(h->allocT(&M->S, 8));
  return AQ_OK;
}

// Solve A p = b to an FP64 relative residual <= tol with MG-preconditioned CG in FP64,
// starting from the current pressure (or zero), and store the result as FP32.
aq_status mgpcg_solve(aq_flip* h, double tol, int max_iters, int check_every, bool from_current,
                      int* iters, double* rel) {
  AQ_TRY(mg_setup(h));
  aq_flip::MgState& M = *h->mg;
  const MgLevel& L0 = M.H.levels[0];
  const long long n = L0.n;
  hipStream_t s = h->stream;
  mg_gather(h->G, h->gr.content, h->gr.rhs, h->gr.p[h->pb], from_current, L0, M.b, M.x, s);
  for (size_t l = 1; l < M.H.levels.size(); ++l) mg_coarsen_flags(M.H.levels[l - 1], M.H.levels[l], s);
  cg_residual(L0, M.b, M.x, M.r, s);
  cg_dot2(n, M.b, M.b, nullptr, nullptr, M.part, M.nparts, M.S, 6, s);   // S[6] = b.b
  mg_vcycle(M.H, M.r, M.z, s);
  cg_copy(n, M.z, M.d, s);
  cg_dot2(n, M.r, M.z, M.r, M.r, M.part, M.nparts, M.S, 0, s);           // S[0] = r.z, S[1] = r.r
  double hs[8] = {0};
  HIP_TRY(hipMemcpyAsync(hs, M.S, sizeof(hs), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  const double bb = hs[6];
  auto relres = [&](double rr) { return bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr); };
  double res = relres(hs[1]);
  int it = 0;
  while (!(res <= tol) && it < max_iters && std::isfinite(res)) {
    cg_apply(L0, M.d, M.q, s);
    cg_dot2(n, M.d, M.q, nullptr, nullptr, M.part, M.nparts, M.S, 2, s);  // S[2] = d.q
    cg_update_xr(n, M.S, M.d, M.q, M.x, M.r, s);
    mg_vcycle(M.H, M.r, M.z, s);
    cg_dot2(n, M.r, M.z, M.r, M.r, M.part, M.nparts, M.S, 4, s);          // S[4] = r.z, S[5] = r.r
    cg_update_d(n, M.S, M.z, M.d, s);
    cg_shift(M.S, s);
    ++it;
    if (it % check_every == 0 || it == max_iters) {
      HIP_TRY(hipMemcpyAsync(hs, M.S, sizeof(hs), hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      res = relres(hs[1]);
    }
  }
  mg_scatter(h->G, L0, M.x, h->gr.p[h->pb], s);
  HIP_TRY(hipGetLastError());
  h->air_clean[h->pb] = true;  // AIR cells written 0
  *iters = it;
  *rel = res;
  return AQ_OK;
}

// Reads back the residual sums the residual kernels left in red_out.
aq_status residual_read(aq_flip* h, double* rel) {
  double o[2];
  HIP_TRY(hipMemcpyAsync(o, h->red_out, sizeof(o), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  *rel = o[1] > 0 ? std::sqrt(o[0]) / std::sqrt(o[1]) : std::sqrt(o[0]);
  return AQ_OK;
}

aq_status residual_now(aq_flip* h, double* rel) {
  launch_residual(h->G, h->gr, h->pb, h->red_partial, h->red_out, h->stream);
  HIP_TRY(hipGetLastError());
  return residual_read(h, rel);
}

// k Jacobi sweeps, then the FP64 residual of the result. Replayed from a captured
// graph (host-launched sweeps cannot keep up with the GPU: about 5.6 against 3.6 us
// per sweep at 128^3); eager while stage timing is on or graphs are switched off.
aq_status sweeps_then_residual(aq_flip* h, int k, double* rel) {
  if (h->timing || !h->use_graphs) {
    jacobi_sweeps(h, k);
    return residual_now(h, rel);
  }
  const int key[4] = {h->pb, h->air_clean[0], h->air_clean[1], k};
  for (auto& g : h->sweep_graphs) {
    if (std::memcmp(g.key, key, sizeof(key)) == 0) {
      HIP_TRY(hipGraphLaunch(g.exec, h->stream));
      h->pb = g.next[0], h->air_clean[0] = g.next[1] != 0, h->air_clean[1] = g.next[2] != 0;
      return residual_read(h, rel);
    }
  }
  if (h->sweep_graphs.size() >= aq_flip::kMaxSweepGraphs) {  // unusual chunk lengths: eager
    jacobi_sweeps(h, k);
    return residual_now(h, rel);
  }
  aq_flip::SweepGraph g{};
  std::memcpy(g.key, key, sizeof(key));
  HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  jacobi_sweeps(h, k);
  launch_residual(h->G, h->gr, h->pb, h->red_partial, h->red_out, h->stream);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(h->stream, &graph);
  if (e != hipSuccess) {
    if (graph) hipGraphDestroy(graph);
    return fail(AQ_EHIP, std::string("sweep capture: ") + hipGetErrorString(e));
  }
  g.graph = graph;
  HIP_TRY(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
  g.next[0] = h->pb, g.next[1] = h->air_clean[0], g.next[2] = h->air_clean[1];
  h->sweep_graphs.push_back(g);
  HIP_TRY(hipGraphLaunch(g.exec, h->stream));
  return residual_read(h, rel);
}

// Jacobi sweeps from the current pressure until the FP64 relative residual is
// <= tol (checked every check_every sweeps) or `budget` sweeps are spent. With
// detect_stall, `stall_checks` consecutive checks without a 1% improvement over
// the best residual (or a non-finite residual) end the solve as stalled.
aq_status sweep_to_tol(aq_flip* h, int budget, bool detect_stall, int* sweeps, double* res,
                       bool* stalled) {
  const aq_flip_solver& S = h->solver;
  *sweeps = 0;
  *stalled = false;
  AQ_TRY(residual_now(h, res));
  double best = *res;
  int flat = 0;
  while (!(*res <= S.tol) && *sweeps < budget) {
    if (!std::isfinite(*res)) {
      *stalled = true;
      break;
    }
    const int k = std::min(S.check_every, budget - *sweeps);
    AQ_TRY(sweeps_then_residual(h, k, res));
    *sweeps += k;
    if (detect_stall) {
      if (std::isfinite(*res) && *res < 0.99 * best) {
        best = *res;
        flat = 0;
      } else if (++flat >= S.stall_checks || !std::isfinite(*res)) {
        *stalled = true;
        break;
      }
    }
  }
  return AQ_OK;
}

// Refinement of a U-Net guess already in the pressure buffer: Jacobi to the solver's
// tol within `budget` sweeps, the network correcting the iterate from its residual
// every nn_refresh sweeps (0: never). While the residual at the end of each chunk of
// sweeps keeps falling by at least 1 % from the previous chunk's, every correction is
// kept: the network removes smooth error that Jacobi is slow on while adding
// high-frequency residual that the following sweeps remove, so the residual right after
// a correction, or even one chunk later on a large grid, says little. Once a chunk
// ends within 1 % of the previous one (Jacobi near its FP32 round-off floor, where the
// network's output is mostly noise), the next correction is judged by the residual
// after the sweeps that follow it: if that is not below the residual before the
// correction, the iterate from before it is restored bit for bit and Jacobi continues
// without further corrections.
aq_status nn_refine(aq_flip* h, aq_unet* net, int budget, int nn_refresh, int* total, double* res,
                    bool* stalled, bool* fallback, int64_t* corrections, int64_t* rejected) {
  const size_t bytes = (size_t)h->G.ncell * sizeof(float);
  int left = budget, sweeps = 0;
  bool correcting = nn_refresh > 0;
  bool judging = false;  // the last chunk followed a correction that is being judged
  double before = 0;     // residual of the iterate saved in p_keep
  double last = 0;       // residual at the end of the previous chunk
  bool have_last = false;
  int pb_keep = 0;
  bool clean_keep = false;
  *total = 0;
  for (;;) {
    const int chunk = correcting ? std::min(nn_refresh, left) : left;
    AQ_TRY(sweep_to_tol(h, chunk, true, &sweeps, res, stalled));
    *total += sweeps;
    left -= sweeps;
    const bool flat = have_last && !(*res < 0.99 * last);
    last = *res;
    have_last = true;
    if (judging) {
      judging = false;
      if (!(*res <= h->solver.tol) && !(*res < before)) {
        // restore the iterate from before the correction
        HIP_TRY(hipMemcpyAsync(h->gr.p[pb_keep], h->p_keep, bytes, hipMemcpyDeviceToDevice, h->stream));
        h->pb = pb_keep;
        h->air_clean[pb_keep] = clean_keep;
        *res = before;
        *stalled = false;
        *rejected += 1;
        correcting = false;
        if (left > 0) continue;
      }
    }
    if (!std::isfinite(*res)) *fallback = true;
    if (*fallback || *res <= h->solver.tol || left <= 0 || *stalled || !correcting) break;
    if (flat) {  // keep a copy of the iterate: this correction will be judged
      if (!h->p_keep) AQ_TRY(h->allocT(&h->p_keep, (size_t)h->G.ncell));
      pb_keep = h->pb;
      clean_keep = h->air_clean[h->pb];
      HIP_TRY(hipMemcpyAsync(h->p_keep, h->gr.p[h->pb], bytes, hipMemcpyDeviceToDevice, h->stream));
    }
    before = *res;
    AQ_TRY(aq_flip_nn_correct(h, net));
    *corrections += 1;
    judging = flat;
  }
  return AQ_OK;
}

// The pressure solve of one step, per the selected mode.
aq_status pressure_stage(aq_flip* h) {
  aq_flip_solver_stats& st = h->sstats;
  if (h->solver.mode == AQ_SOLVE_FIXED) {
    h->begin(AQ_STAGE_PRESSURE);  // the Jacobi sweeps
    fixed_sweeps(h);
    h->end(AQ_STAGE_PRESSURE, h->cfg.pressure_iters);
    return AQ_OK;
  }
  h->begin(AQ_STAGE_PRESSURE);
  int sweeps = 0, total = 0;
  double res = 0, guess = 1.0;
  bool stalled = false, fallback = false;
  bool mg_converged = false;  // MGPCG: the CG residual reached tol
  if (h->solver.mode == AQ_SOLVE_TOL_NN) {
    if (!h->net) return fail(AQ_EINVAL, "AQ_SOLVE_TOL_NN needs a U-Net");
    AQ_TRY(aq_flip_nn_guess(h, h->net));
    h->air_clean[h->pb] = true;  // the guess is 0 on AIR cells
    AQ_TRY(residual_now(h, &guess));
    // Jacobi converges to the same fixed point from any finite start, so only a
    // non-finite guess or refinement is rejected for the zero-guess solve. A finite
    // guess that stalls has reached FP32 Jacobi's round-off floor, where a solve from
    // zero would stall as well.
    if (!std::isfinite(guess)) fallback = true;
    // Refine with Jacobi; every nn_refresh sweeps (if set) the network corrects the
    // iterate from its residual, then Jacobi continues.
    if (!fallback)
      AQ_TRY(nn_refine(h, h->net, h->solver.max_iters, h->solver.nn_refresh, &total, &res, &stalled,
                       &fallback, &st.nn_corrections, &st.nn_rejected));
  }
  if (h->solver.mode == AQ_SOLVE_MGPCG) {
    const bool from_current = h->cfg.pressure_init == AQ_PRESSURE_PREVIOUS;
    int it = 0;
    double cg = 0;
    AQ_TRY(mgpcg_solve(h, h->solver.tol, h->solver.max_iters, h->solver.check_every, from_current, &it, &cg));
    AQ_TRY(residual_now(h, &res));  // of the stored FP32 field (reported)
    mg_converged = cg <= h->solver.tol;
    total = it;
    guess = 1.0;
  }
  if (h->solver.mode == AQ_SOLVE_TOL_ZERO || fallback) {
    zero_pressure(h);
    AQ_TRY(sweep_to_tol(h, std::max(0, h->solver.max_iters), true, &sweeps, &res, &stalled));
    total += sweeps;
    if (h->solver.mode == AQ_SOLVE_TOL_ZERO) guess = 1.0;
  }
  h->end(AQ_STAGE_PRESSURE, total);
  st.solves += 1;
  st.sweeps += total;
  st.fallbacks += fallback ? 1 : 0;
  const bool reached = h->solver.mode == AQ_SOLVE_MGPCG ? mg_converged : res <= h->solver.tol;
  st.capped += reached ? 0 : 1;
  st.last_sweeps = total;
  st.last_residual = res;
  st.last_guess_residual = guess;
  st.last_fallback = fallback ? 1 : 0;
  return AQ_OK;
}

// Fluid::simulationStep (Fluid/Fluid_3D_FLIP.cu:316-352). G2P, advection, hashing
// and bin counting are one particle pass; gravity and boundary one face pass.
// The bin counts are zero on entry: allocated so, and every scan sets them back to zero.
aq_status enqueue_step(aq_flip* h) {
  h->begin(AQ_STAGE_G2P);
  launch_g2p_advect_hash(h->G, h->P[h->cur], h->Np, h->K, h->gr, h->nv, h->nf, h->sp, h->B, h->stream);
  h->end(AQ_STAGE_G2P);
  h->begin(AQ_STAGE_TRANSFER_TO_GRID);
  transfer_after_count(h, true);  // ends the bucket before P2G, which is timed as AQ_STAGE_P2G
  h->begin(AQ_STAGE_GRAVITY);
  launch_gravity_boundary(h->G, h->gr, h->nv, h->sp, 1, 1, h->stream);
  h->end(AQ_STAGE_GRAVITY);
  h->begin(AQ_STAGE_DIVERGENCE);
  launch_divergence(h->G, h->gr, h->nv, h->B, h->stream);
  h->end(AQ_STAGE_DIVERGENCE);
  // Diffusion reads the fractions, content and active-block flags, all final after
  // the divergence pass, and writes only the fractions, which nothing before the
  // next step's G2P reads. (Stage timing keeps the reference's sequential order.)
  const bool fused = h->fuse && h->solver.mode == AQ_SOLVE_FIXED &&
                     h->cfg.diffusion_iters > 0 && h->K > 1 && h->cfg.pressure_iters > 0;
  const bool side = !fused && h->overlap && !h->timing && h->cfg.diffusion_iters > 0 && h->K > 1;
  if (side) {
    HIP_TRY(hipEventRecord(h->ev_fork, h->stream));
    HIP_TRY(hipStreamWaitEvent(h->side, h->ev_fork, 0));
    diffusion_sweeps(h, h->side);
    HIP_TRY(hipEventRecord(h->ev_join, h->side));
  }
  if (fused) {
    fused_pressure_diffusion(h);
  } else if (const aq_status ps = pressure_stage(h); ps != AQ_OK) {
    if (side) hipStreamWaitEvent(h->stream, h->ev_join, 0);  // never leave the fork open
    return ps;
  }
  h->begin(AQ_STAGE_PROJECT);
  launch_project(h->G, h->gr, h->nv, h->hv, h->pb, h->stream);
  h->end(AQ_STAGE_PROJECT);
  h->begin(AQ_STAGE_EXTRAPOLATE);
  stage_extrapolate(h, true);
  h->end(AQ_STAGE_EXTRAPOLATE, 3);
  if (side) {
    if (h->cfg.pressure_iters == 0) launch_air_pressure_zero(h->G, h->gr, h->pb, h->stream);
    HIP_TRY(hipStreamWaitEvent(h->stream, h->ev_join, 0));
  } else if (!fused) {
    h->begin(AQ_STAGE_DIFFUSION);
    h->end(AQ_STAGE_DIFFUSION, stage_diffusion(h, true));
  }
  HIP_TRY(hipGetLastError());
  h->physicalTime += h->cfg.dt;
  return AQ_OK;
}

void drop_graphs(aq_flip* h) {
  hipStreamSynchronize(h->stream);
  for (auto& g : h->graphs) {
    hipGraphExecDestroy(g.exec);
    hipGraphDestroy(g.graph);
  }
  h->graphs.clear();
}

void state_of(const aq_flip* h, int k[7]) {
  k[0] = h->cur, k[1] = h->nv, k[2] = h->hv, k[3] = h->pb, k[4] = h->nf;
  k[5] = h->air_clean[0], k[6] = h->air_clean[1];
}
void set_state(aq_flip* h, const int k[7]) {
  h->cur = k[0], h->nv = k[1], h->hv = k[2], h->pb = k[3], h->nf = k[4];
  h->air_clean[0] = k[5] != 0, h->air_clean[1] = k[6] != 0;
}

// One step through a captured graph (eager launches when timing is on, so the
// stage events sit on the stream between kernels).
// A step's G2P reads new and old velocity from gr.nvo, which the previous step's last
// extrapolation pass wrote on every row a particle's first sample can touch (the rows
// within 4 of a cell row holding FLUID; a particle's corners lie within one row of its
// FLUID cell). Anything else that changes the velocities, or a fresh handle, leaves it
// stale, and the copy is then rebuilt in full first.
aq_status graph_step(aq_flip* h) {
  if (!h->nvo_valid) {
    launch_build_nvo(h->G, h->gr, h->nv, h->stream);
    h->nvo_valid = true;
  }
  if (h->timing || !h->use_graphs || h->solver.mode != AQ_SOLVE_FIXED) return enqueue_step(h);
  int key[7];
  state_of(h, key);
  for (auto& g : h->graphs) {
    if (std::memcmp(g.key, key, sizeof(key)) == 0) {
      HIP_TRY(hipGraphLaunch(g.exec, h->stream));
      set_state(h, g.next);
      h->physicalTime += h->cfg.dt;
      return AQ_OK;
    }
  }
  aq_flip::StepGraph g{};
  std::memcpy(g.key, key, sizeof(key));
  HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  aq_status st = enqueue_step(h);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(h->stream, &graph);
  if (st != AQ_OK) {
    if (graph) hipGraphDestroy(graph);
    return st;
  }
  if (e != hipSuccess) return fail(AQ_EHIP, std::string("step capture: ") + hipGetErrorString(e));
  g.graph = graph;
  HIP_TRY(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
  state_of(h, g.next);
  h->graphs.push_back(g);
  HIP_TRY(hipGraphLaunch(g.exec, h->stream));
  return AQ_OK;
}

// ---- particle seeding on the GPU ------------------------------------------
// The volumes of a scene as the GPU seeder walks them, with per-block output
// offsets from a counting pass (seed_kernels.hip); identical to seed_scene's output.
struct SeedPlan {
  SeedParams sp{};
  std::vector<SeedVolume> vols;
  std::vector<long long> nc;                // candidates per volume (cells x 8)
  std::vector<std::vector<int>> offsets;    // per volume: exclusive prefix of block counts
  std::vector<long long> first;             // per volume: index of its first particle
  long long total = 0;
};

bool in_unit_cube(const float c[3]) {
  for (int a = 0; a < 3; ++a)
    if (c[a] < 0 || c[a] > 1) return false;
  return true;
}

aq_status plan_seeding(const aq_flip_config& cfg, SeedPlan* plan) {
  const int S[3] = {cfg.size[0], cfg.size[1], cfg.size[2]};
  const float h = cfg.grid_physical_size / (float)cfg.size[1];
  plan->sp.seed = cfg.seed;
  plan->sp.h = h;
  plan->sp.jitter = cfg.particle_jitter;
  for (int a = 0; a < 3; ++a) plan->sp.S[a] = S[a];
  uint64_t draws = 0;
  for (int i = 0; i < cfg.n_volumes; ++i) {
    const aq_flip_volume& v = cfg.volumes[i];
    SeedVolume sv{};
    sv.phase = v.phase;
    if (v.shape == AQ_SHAPE_BOX) {
      if (!in_unit_cube(v.box_min) || !in_unit_cube(v.box_max)) continue;
      // createSquareFluid's loops: z, then x, then y, each from (int)(min*S) while (float)i < max*S
      const int order[3] = {2, 0, 1};
      for (int j = 0; j < 3; ++j) {
        const int a = order[j];
        const int lo = (int)(v.box_min[a] * (float)S[a]);
        int cnt = 0;
        for (int c = lo; (float)c < v.box_max[a] * (float)S[a]; ++c) ++cnt;
        sv.n[j] = cnt;
        sv.o[j] = lo;
        sv.axis[j] = a;
      }
    } else if (v.shape == AQ_SHAPE_SPHERE) {
      if (!in_unit_cube(v.center)) continue;
      // createSphereFluid visits every cell: y, then x, then z
      const int order[3] = {1, 0, 2};
      for (int j = 0; j < 3; ++j) {
        sv.n[j] = S[order[j]];
        sv.o[j] = 0;
        sv.axis[j] = order[j];
      }
      sv.sphere = 1;
      for (int a = 0; a < 3; ++a) sv.cc[a] = (v.center[a] * (float)S[a]) * h;
      sv.limit = (v.radius * h) * (float)S[1];
    } else {
      continue;
    }
    sv.base = draws;
    const long long cells = (long long)std::max(sv.n[0], 0) * std::max(sv.n[1], 0) * std::max(sv.n[2], 0);
    draws += (uint64_t)cells * 24u;
    if (cells == 0) continue;
    plan->vols.push_back(sv);
    plan->nc.push_back(cells * 8);
  }
  if (plan->vols.empty()) return AQ_OK;  // nothing to seed (no device work)
  // counting pass
  hipStream_t s = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  const int B = seed_block_size();
  long long maxb = 1;
  for (long long nc : plan->nc) maxb = std::max(maxb, (nc + B - 1) / B);
  int* d_counts = nullptr;
  hipError_t e = hipMalloc((void**)&d_counts, (size_t)maxb * sizeof(int));
  for (size_t i = 0; i < plan->vols.size() && e == hipSuccess; ++i) {
    const long long nb = (plan->nc[i] + B - 1) / B;
    launch_seed_count(plan->vols[i], plan->sp, plan->nc[i], d_counts, s);
    std::vector<int> counts((size_t)nb);
    e = hipMemcpyAsync(counts.data(), d_counts, (size_t)nb * sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    std::vector<int> off((size_t)nb);
    long long run = 0;
    for (long long b = 0; b < nb; ++b) {
      off[(size_t)b] = (int)run;
      run += counts[(size_t)b];
    }
    plan->first.push_back(plan->total);
    plan->total += run;
    plan->offsets.push_back(std::move(off));
  }
  if (d_counts) hipFree(d_counts);
  hipStreamDestroy(s);
  if (e != hipSuccess) return fail(AQ_EHIP, std::string("seeding: ") + hipGetErrorString(e));
  if (plan->total > (1ll << 31) - 1) return fail(AQ_EINVAL, "too many particles");
  return AQ_OK;
}

// Write the planned particles into the handle's current particle buffer.
aq_status write_seeded(aq_flip* h, const SeedPlan& plan) {
  long long maxb = 1;
  for (const auto& o : plan.offsets) maxb = std::max<long long>(maxb, (long long)o.size());
  int* d_off = nullptr;
  HIP_TRY(hipMalloc((void**)&d_off, (size_t)maxb * sizeof(int)));
  hipError_t e = hipSuccess;
  for (size_t i = 0; i < plan.vols.size() && e == hipSuccess; ++i) {
    e = hipMemcpyAsync(d_off, plan.offsets[i].data(), plan.offsets[i].size() * sizeof(int),
                       hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
      launch_seed_write(plan.vols[i], plan.sp, plan.nc[i], d_off, plan.first[i], h->P[h->cur], h->stream);
      e = hipStreamSynchronize(h->stream);  // d_off is reused by the next volume
    }
  }
  hipFree(d_off);
  if (e != hipSuccess) return fail(AQ_EHIP, std::string("seeding: ") + hipGetErrorString(e));
  return AQ_OK;
}

// Switches for measurements: AQFLIP_NO_GRAPH=1 (eager launches), AQFLIP_NO_OVERLAP=1
// (diffusion after the pressure solve on one stream), AQFLIP_NO_FUSE=1 (no fused
// pressure + diffusion sweeps), AQFLIP_NO_FIXED_POINT_SKIP=1 (paired diffusion sweeps
// always computed). Results are identical in every combination.
void read_env(aq_flip* h) {
  if (const char* env = std::getenv("AQFLIP_NO_GRAPH")) h->use_graphs = env[0] == '0';
  if (const char* env = std::getenv("AQFLIP_NO_OVERLAP")) h->overlap = env[0] != '1';
  if (const char* env = std::getenv("AQFLIP_NO_FUSE")) h->fuse = env[0] != '1';
  if (const char* env = std::getenv("AQFLIP_NO_FIXED_POINT_SKIP")) h->skip_fixed = env[0] != '1';
  if (const char* env = std::getenv("AQFLIP_LISTED_GRID")) h->listed_grid = std::max(0, std::atoi(env));
}

aq_status create_common(const aq_flip_config* cfg, const float* pos, const float* vel,
                        const float* frac4, long long n, aq_flip** out, const SeedPlan* plan = nullptr) {
  aq_flip* h = new aq_flip;
  read_env(h);
  aq_status st = setup(h, cfg, n);
  if (plan) {  // seeded on the device: positions are clamped into the domain by construction
    h->K = channels_for(*cfg, nullptr, 0);
    if (st == AQ_OK) st = write_seeded(h, *plan);
  } else {
    if (st == AQ_OK) st = check_positions(h, pos, n);
    if (st == AQ_OK) {
      h->K = channels_for(*cfg, frac4, n);
      st = upload_particles(h, pos, vel, frac4);
    }
  }
  if (st == AQ_OK) {
    // Fluid::init: fixBoundary, then transferToGrid (Fluid_3D_FLIP.cu:485,500).
    launch_gravity_boundary(h->G, h->gr, h->nv, h->sp, 0, 1, h->stream);
    stage_transfer(h);
    refresh_rhs(h);
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) st = fail(AQ_EHIP, std::string("init: ") + hipGetErrorString(e));
  }
  if (st != AQ_OK) {
    delete h;
    return st;
  }
  *out = h;
  return AQ_OK;
}

size_t field_elems(const aq_flip* h, int field, int* width, bool* isInt) {
  const Geom& G = h->G;
  const size_t e1 = (size_t)(G.X + 1) * (G.Y + 1) * (G.Z + 1), e0 = (size_t)G.X * G.Y * G.Z;
  *isInt = false;
  switch (field) {
    case AQ_FIELD_CONTENT: *width = 1; *isInt = true; return e1;
    case AQ_FIELD_VELOCITY:
    case AQ_FIELD_NEW_VELOCITY: *width = 4; return e1;
    case AQ_FIELD_PRESSURE:
    case AQ_FIELD_DIVERGENCE:
    case AQ_FIELD_DENSITY: *width = 1; return e0;
    case AQ_FIELD_PARTICLE_COUNT: *width = 1; *isInt = true; return e0;
    case AQ_FIELD_HAS_VELOCITY: *width = 4; *isInt = true; return e0;
    case AQ_FIELD_VOLUME_FRACTIONS:
    case AQ_FIELD_NEW_VOLUME_FRACTIONS: *width = 4; return e0;
  }
  *width = 0;
  return 0;
}

template <typename T>
aq_status d2h(aq_flip* h, std::vector<T>& dst, const void* src, size_t n) {
  dst.resize(n);
  HIP_TRY(hipMemcpyAsync(dst.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return AQ_OK;
}
template <typename T>
aq_status h2d(aq_flip* h, void* dst, const std::vector<T>& src) {
  HIP_TRY(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return AQ_OK;
}

}  // namespace

namespace aqf {
FlipView flip_view(aq_flip* h) { return FlipView{h->G, &h->gr, &h->pb, h->stream}; }
void set_last_error(const std::string& msg) { g_err = msg; }
}  // namespace aqf

extern "C" {

aq_status aq_flip_set_solver(aq_flip* h, const aq_flip_solver* s, aq_unet* net) {
  if (!h || !s) return fail(AQ_EINVAL, "null argument");
  if (s->mode < AQ_SOLVE_FIXED || s->mode > AQ_SOLVE_MGPCG) return fail(AQ_EINVAL, "unknown solver mode");
  if (s->mode != AQ_SOLVE_FIXED && (!(s->tol > 0) || s->max_iters < 0 || s->check_every < 1 ||
                                    s->stall_checks < 1 || s->nn_refresh < 0))
    return fail(AQ_EINVAL, "tolerance solves need tol > 0, max_iters >= 0, check_every >= 1, "
                           "stall_checks >= 1, nn_refresh >= 0");
  if (s->mode == AQ_SOLVE_TOL_NN && !net) return fail(AQ_EINVAL, "AQ_SOLVE_TOL_NN needs a U-Net");
  if (s->mode == AQ_SOLVE_TOL_NN && (h->G.X % 4 || h->G.Y % 4 || h->G.Z % 4))
    return fail(AQ_EINVAL, "the U-Net needs grid extents that are multiples of 4");
  h->solver = *s;
  h->net = net;
  return AQ_OK;
}

aq_status aq_flip_solve_pressure_nn(aq_flip* h, aq_unet* net, double tol, int32_t max_iters,
                                    int32_t check_every, int32_t nn_refresh, int32_t* iters_out,
                                    double* rel_out, int32_t* corrections_out, int32_t* rejected_out) {
  if (!h || !net || max_iters < 0 || check_every < 1 || nn_refresh < 0) return fail(AQ_EINVAL, "bad argument");
  const aq_flip_solver keep = h->solver;
  h->solver.tol = tol;
  h->solver.check_every = check_every;
  int total = 0;
  double res = 0;
  bool stalled = false, fallback = false;
  int64_t corr = 0, rej = 0;
  aq_status st = aq_flip_nn_guess(h, net);
  if (st == AQ_OK) {
    h->air_clean[h->pb] = true;  // the guess is 0 on AIR cells
    st = nn_refine(h, net, max_iters, nn_refresh, &total, &res, &stalled, &fallback, &corr, &rej);
  }
  h->solver = keep;
  if (st != AQ_OK) return st;
  if (iters_out) *iters_out = total;
  if (rel_out) *rel_out = res;
  if (corrections_out) *corrections_out = (int32_t)corr;
  if (rejected_out) *rejected_out = (int32_t)rej;
  if (!(res <= tol)) return fail(AQ_ESTALL, fallback ? "U-Net refinement is not finite" : "pressure solve did not reach tol");
  return AQ_OK;
}

aq_status aq_flip_solver_stats_get(aq_flip* h, aq_flip_solver_stats* out, int32_t reset) {
  if (!h || !out) return fail(AQ_EINVAL, "null argument");
  *out = h->sstats;
  if (reset) h->sstats = aq_flip_solver_stats{};
  return AQ_OK;
}

aq_status aq_flip_config_default(aq_flip_config* c) {
  if (!c) return fail(AQ_EINVAL, "null config");
  std::memset(c, 0, sizeof(*c));
  // ConfigFLIP (Fluid/FluidConfig.cuh:41-50)
  c->size[0] = c->size[1] = c->size[2] = 50;
  c->pressure_iters = 100;
  c->diffusion_iters = 100;
  c->dt = 0.033f;
  c->flip_coeff = 0.95f;
  c->particle_jitter = 0.05f;
  c->grid_physical_size = 10.f;  // Fluid_3D_FLIP.cuh:62
  // FluidConfig (Fluid/FluidConfig.cuh:68-97)
  c->gravity[0] = 0.f;
  c->gravity[1] = -9.8f;
  c->gravity[2] = 0.f;
  c->phase_count = 2;
  for (int a = 0; a < 4; ++a) c->phase_densities[a] = 1.f;
  c->diffusion_coeff = 0.001f;
  c->pressure_init = AQ_PRESSURE_PREVIOUS;
  c->ref_boundary_coverage = 1;
  c->seed = 0;
  // Default scene: a pool and a ball (InitializationVolume, FluidConfig.cuh:20-39).
  c->n_volumes = 2;
  aq_flip_volume& box = c->volumes[0];
  box.shape = AQ_SHAPE_BOX;
  box.box_min[0] = box.box_min[1] = box.box_min[2] = 0.f;
  box.box_max[0] = 1.f;
  box.box_max[1] = 0.1875f;
  box.box_max[2] = 1.f;
  box.center[0] = 0.5f;
  box.center[1] = 0.75f;
  box.center[2] = 0.5f;
  box.radius = 0.125f;
  box.phase = 0;
  c->volumes[1] = box;
  c->volumes[1].shape = AQ_SHAPE_SPHERE;
  return AQ_OK;
}

aq_status aq_flip_seed_particles(const aq_flip_config* cfg, float* pos, float* frac4,
                                 int64_t capacity, int64_t* n_out) {
  AQ_TRY(validate(cfg));
  if (!n_out) return fail(AQ_EINVAL, "null n_out");
  if (!pos) {
    *n_out = seed_scene(*cfg, nullptr, nullptr);
    return AQ_OK;
  }
  std::vector<float> p, f;
  const int64_t n = seed_scene(*cfg, &p, &f);
  *n_out = n;
  if (capacity < n) return fail(AQ_EINVAL, "capacity smaller than the particle count");
  std::memcpy(pos, p.data(), p.size() * sizeof(float));
  if (frac4) std::memcpy(frac4, f.data(), f.size() * sizeof(float));
  return AQ_OK;
}

aq_status aq_flip_create(const aq_flip_config* cfg, aq_flip** out) {
  AQ_TRY(validate(cfg));
  if (!out) return fail(AQ_EINVAL, "null out");
  SeedPlan plan;
  AQ_TRY(plan_seeding(*cfg, &plan));
  if (plan.total <= 0) return fail(AQ_EINVAL, "the configured volumes produce no particles");
  return create_common(cfg, nullptr, nullptr, nullptr, plan.total, out, &plan);
}

aq_status aq_flip_create_with_particles(const aq_flip_config* cfg, const float* pos,
                                        const float* vel, const float* frac4, int64_t n,
                                        aq_flip** out) {
  AQ_TRY(validate(cfg));
  if (!out || !pos) return fail(AQ_EINVAL, "null argument");
  if (n <= 0 || n > (1ll << 31) - 1) return fail(AQ_EINVAL, "particle count out of range");
  return create_common(cfg, pos, vel, frac4, n, out);
}

void aq_flip_destroy(aq_flip* h) { delete h; }

aq_status aq_flip_step(aq_flip* h) {
  if (!h) return fail(AQ_EINVAL, "null handle");
  AQ_TRY(graph_step(h));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return h->collect();
}

aq_status aq_flip_steps(aq_flip* h, int32_t count) {
  if (!h || count < 0) return fail(AQ_EINVAL, "bad argument");
  for (int i = 0; i < count; ++i) AQ_TRY(graph_step(h));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return h->collect();
}

aq_status aq_flip_run_stage(aq_flip* h, int32_t stage) {
  if (!h) return fail(AQ_EINVAL, "null handle");
  if (stage < 0 || stage >= AQ_STAGE_COUNT) return fail(AQ_EINVAL, "unknown stage");
  AQ_TRY(run_stage(h, stage));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return h->collect();
}

int64_t aq_flip_particle_count(const aq_flip* h) { return h ? h->Np : 0; }

aq_status aq_flip_get_particles(aq_flip* h, float* pos, float* vel, float* frac4) {
  if (!h) return fail(AQ_EINVAL, "null handle");
  const long long n = h->Np;
  std::vector<float4> a, b, c;
  const Particles& P = h->P[h->cur];
  AQ_TRY(d2h(h, a, P.a, (size_t)n));
  AQ_TRY(d2h(h, b, P.b, (size_t)n));
  if (h->K > 2) AQ_TRY(d2h(h, c, P.c, (size_t)n));
  for (long long i = 0; i < n; ++i) {
    if (pos) {
      pos[3 * i] = a[i].x;
      pos[3 * i + 1] = a[i].y;
      pos[3 * i + 2] = a[i].z;
    }
    if (vel) {
      vel[3 * i] = a[i].w;
      vel[3 * i + 1] = b[i].x;
      vel[3 * i + 2] = b[i].y;
    }
    if (frac4) {
      frac4[4 * i] = b[i].z;
      frac4[4 * i + 1] = h->K > 1 ? b[i].w : 0.f;
      frac4[4 * i + 2] = h->K > 2 ? c[i].x : 0.f;
      frac4[4 * i + 3] = h->K > 3 ? c[i].y : 0.f;
    }
  }
  return AQ_OK;
}

aq_status aq_flip_set_particles(aq_flip* h, const float* pos, const float* vel, const float* frac4,
                                int64_t n) {
  if (!h || !pos) return fail(AQ_EINVAL, "null argument");
  if (n != h->Np) return fail(AQ_EINVAL, "particle count differs from the handle's");
  AQ_TRY(check_positions(h, pos, n));
  const int K = std::max(h->K, channels_for(h->cfg, frac4, n));
  if (K != h->K) drop_graphs(h);
  h->K = K;
  return upload_particles(h, pos, vel, frac4);
}

size_t aq_flip_field_bytes(const aq_flip* h, int32_t field) {
  if (!h) return 0;
  int w;
  bool isInt;
  const size_t n = field_elems(h, field, &w, &isInt);
  return n * (size_t)w * 4;
}

aq_status aq_flip_get_field(aq_flip* h, int32_t field, void* dst, size_t bytes) {
  if (!h || !dst) return fail(AQ_EINVAL, "null argument");
  int w;
  bool isInt;
  const size_t n = field_elems(h, field, &w, &isInt);
  if (n == 0) return fail(AQ_EINVAL, "unknown field");
  if (bytes != n * w * 4) return fail(AQ_EINVAL, "byte count does not match the field");
  const Geom& G = h->G;
  const size_t nc = (size_t)G.ncell;
  float* F = (float*)dst;
  int32_t* I = (int32_t*)dst;
  const bool face = (field == AQ_FIELD_CONTENT || field == AQ_FIELD_VELOCITY ||
                     field == AQ_FIELD_NEW_VELOCITY);
  const int EX = face ? G.X + 1 : G.X, EYe = face ? G.Y + 1 : G.Y, EZe = face ? G.Z + 1 : G.Z;
  auto each = [&](auto&& fn) {
    size_t o = 0;
    for (int x = 0; x < EX; ++x)
      for (int y = 0; y < EYe; ++y)
        for (int z = 0; z < EZe; ++z, ++o) fn(o, (size_t)G.g(x, y, z));
  };
  switch (field) {
    case AQ_FIELD_CONTENT: {
      std::vector<uint8_t> c;
      AQ_TRY(d2h(h, c, h->gr.content, nc));
      each([&](size_t o, size_t g) { I[o] = c[g]; });
      break;
    }
    case AQ_FIELD_VELOCITY:
    case AQ_FIELD_NEW_VELOCITY: {
      std::vector<float> comp[3];
      for (int a = 0; a < 3; ++a)
        AQ_TRY(d2h(h, comp[a], field == AQ_FIELD_VELOCITY ? h->gr.vel[a] : h->gr.nvel[h->nv][a], nc));
      each([&](size_t o, size_t g) {
        for (int a = 0; a < 3; ++a) F[4 * o + a] = comp[a][g];
        F[4 * o + 3] = 0.f;
      });
      break;
    }
    case AQ_FIELD_PRESSURE:
    case AQ_FIELD_DIVERGENCE:
    case AQ_FIELD_DENSITY: {
      const float* src = field == AQ_FIELD_PRESSURE ? h->gr.p[h->pb]
                         : field == AQ_FIELD_DIVERGENCE ? h->gr.div
                                                        : h->gr.density;
      std::vector<float> v;
      AQ_TRY(d2h(h, v, src, nc));
      each([&](size_t o, size_t g) { F[o] = v[g]; });
      break;
    }
    case AQ_FIELD_PARTICLE_COUNT: {
      std::vector<int> off;
      AQ_TRY(d2h(h, off, h->B.offsets, (size_t)G.ncells_hash + 1));
      for (int c = 0; c < G.ncells_hash; ++c) {
        const int cnt = off[c + 1] - off[c];
        I[c] = cnt > 0 ? cnt - 1 : 0;  // cellEnd - cellStart (Fluid_3D_FLIP.cu:180)
      }
      break;
    }
    case AQ_FIELD_HAS_VELOCITY: {
      std::vector<uint8_t> v;
      AQ_TRY(d2h(h, v, h->gr.hasvel[h->hv], nc));
      each([&](size_t o, size_t g) {
        for (int a = 0; a < 3; ++a) I[4 * o + a] = (v[g] >> a) & 1;
        I[4 * o + 3] = 0;
      });
      break;
    }
    case AQ_FIELD_VOLUME_FRACTIONS:
    case AQ_FIELD_NEW_VOLUME_FRACTIONS: {
      std::vector<float> comp[4];
      for (int a = 0; a < h->K; ++a)
        AQ_TRY(d2h(h, comp[a], field == AQ_FIELD_VOLUME_FRACTIONS ? h->gr.vf[a] : h->gr.nvf[h->nf][a], nc));
      each([&](size_t o, size_t g) {
        for (int a = 0; a < 4; ++a) F[4 * o + a] = a < h->K ? comp[a][g] : 0.f;
      });
      break;
    }
  }
  return AQ_OK;
}

aq_status aq_flip_set_field(aq_flip* h, int32_t field, const void* src, size_t bytes) {
  if (!h || !src) return fail(AQ_EINVAL, "null argument");
  h->nvo_valid = false;
  int w;
  bool isInt;
  const size_t n = field_elems(h, field, &w, &isInt);
  if (n == 0) return fail(AQ_EINVAL, "unknown field");
  if (field == AQ_FIELD_PARTICLE_COUNT)
    return fail(AQ_EINVAL, "particleCount is derived from the cell bins and cannot be set");
  if (bytes != n * w * 4) return fail(AQ_EINVAL, "byte count does not match the field");
  const Geom& G = h->G;
  const size_t nc = (size_t)G.ncell;
  const float* F = (const float*)src;
  const int32_t* I = (const int32_t*)src;
  const bool face = (field == AQ_FIELD_CONTENT || field == AQ_FIELD_VELOCITY ||
                     field == AQ_FIELD_NEW_VELOCITY);
  const int EX = face ? G.X + 1 : G.X, EYe = face ? G.Y + 1 : G.Y, EZe = face ? G.Z + 1 : G.Z;
  auto each = [&](auto&& fn) {
    size_t o = 0;
    for (int x = 0; x < EX; ++x)
      for (int y = 0; y < EYe; ++y)
        for (int z = 0; z < EZe; ++z, ++o) fn(o, (size_t)G.g(x, y, z));
  };
  switch (field) {
    case AQ_FIELD_CONTENT: {
      std::vector<uint8_t> c;
      AQ_TRY(d2h(h, c, h->gr.content, nc));
      each([&](size_t o, size_t g) { c[g] = (uint8_t)I[o]; });
      AQ_TRY(h2d(h, h->gr.content, c));
      content_changed(h);
      refresh_rhs(h);
      HIP_TRY(hipStreamSynchronize(h->stream));
      break;
    }
    case AQ_FIELD_VELOCITY:
    case AQ_FIELD_NEW_VELOCITY: {
      for (int a = 0; a < 3; ++a) {
        float* d = field == AQ_FIELD_VELOCITY ? h->gr.vel[a] : h->gr.nvel[h->nv][a];
        std::vector<float> v;
        AQ_TRY(d2h(h, v, d, nc));
        each([&](size_t o, size_t g) { v[g] = F[4 * o + a]; });
        AQ_TRY(h2d(h, d, v));
      }
      break;
    }
    case AQ_FIELD_PRESSURE:
    case AQ_FIELD_DIVERGENCE:
    case AQ_FIELD_DENSITY: {
      float* d = field == AQ_FIELD_PRESSURE ? h->gr.p[h->pb]
                 : field == AQ_FIELD_DIVERGENCE ? h->gr.div
                                                : h->gr.density;
      std::vector<float> v;
      AQ_TRY(d2h(h, v, d, nc));
      each([&](size_t o, size_t g) { v[g] = F[o]; });
      AQ_TRY(h2d(h, d, v));
      if (field == AQ_FIELD_PRESSURE) h->air_clean[h->pb] = false;
      if (field != AQ_FIELD_PRESSURE) {
        // keep the Jacobi right-hand side consistent with divergence*density
        refresh_rhs(h);
        HIP_TRY(hipStreamSynchronize(h->stream));
      }
      break;
    }
    case AQ_FIELD_HAS_VELOCITY: {
      std::vector<uint8_t> v;
      AQ_TRY(d2h(h, v, h->gr.hasvel[h->hv], nc));
      each([&](size_t o, size_t g) {
        v[g] = (uint8_t)((I[4 * o] ? 1 : 0) | (I[4 * o + 1] ? 2 : 0) | (I[4 * o + 2] ? 4 : 0));
      });
      AQ_TRY(h2d(h, h->gr.hasvel[h->hv], v));
      break;
    }
    case AQ_FIELD_VOLUME_FRACTIONS:
    case AQ_FIELD_NEW_VOLUME_FRACTIONS: {
      int need = 1;
      for (size_t o = 0; o < n; ++o)
        for (int a = 3; a >= need; --a)
          if (F[4 * o + a] != 0.f) {
            need = a + 1;
            break;
          }
      if (need > h->K) {
        drop_graphs(h);
        h->K = need;
      }
      for (int a = 0; a < h->K; ++a) {
        float* d = field == AQ_FIELD_VOLUME_FRACTIONS ? h->gr.vf[a] : h->gr.nvf[h->nf][a];
        std::vector<float> v;
        AQ_TRY(d2h(h, v, d, nc));
        each([&](size_t o, size_t g) { v[g] = F[4 * o + a]; });
        AQ_TRY(h2d(h, d, v));
      }
      break;
    }
  }
  return AQ_OK;
}

aq_status aq_flip_get_cells(aq_flip* h, int32_t* cell_start, int32_t* cell_end, int32_t* hash,
                            int32_t* perm) {
  if (!h) return fail(AQ_EINVAL, "null handle");
  const Geom& G = h->G;
  std::vector<int> off;
  AQ_TRY(d2h(h, off, h->B.offsets, (size_t)G.ncells_hash + 1));
  const size_t e1 = (size_t)(G.X + 1) * (G.Y + 1) * (G.Z + 1);
  if (cell_start || cell_end) {
    // Reference convention: cellStart/cellEnd sized (X+1)(Y+1)(Z+1), inclusive end,
    // -1 where empty (Fluid_3D_common.cuh:47-48, Fluid_kernels.cuh:3-22).
    for (size_t c = 0; c < e1; ++c) {
      int s = -1, e = -1;
      if (c < (size_t)G.ncells_hash && off[c + 1] > off[c]) {
        s = off[c];
        e = off[c + 1] - 1;
      }
      if (cell_start) cell_start[c] = s;
      if (cell_end) cell_end[c] = e;
    }
  }
  if (hash)
    for (int c = 0; c < G.ncells_hash; ++c)
      for (int k = off[c]; k < off[c + 1]; ++k) hash[k] = c;
  if (perm) {
    std::vector<int> p;
    AQ_TRY(d2h(h, p, h->B.perm, (size_t)h->Np));
    std::memcpy(perm, p.data(), p.size() * sizeof(int));
  }
  return AQ_OK;
}

aq_status aq_flip_residual_fp64(aq_flip* h, double* rel) {
  if (!h || !rel) return fail(AQ_EINVAL, "null argument");
  launch_residual(h->G, h->gr, h->pb, h->red_partial, h->red_out, h->stream);
  HIP_TRY(hipGetLastError());
  double o[2];
  HIP_TRY(hipMemcpyAsync(o, h->red_out, sizeof(o), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  *rel = o[1] > 0 ? std::sqrt(o[0]) / std::sqrt(o[1]) : std::sqrt(o[0]);
  return AQ_OK;
}

aq_status aq_flip_volume_stats(aq_flip* h, aq_flip_stats* out) {
  if (!h || !out) return fail(AQ_EINVAL, "null argument");
  launch_stats(h->G, h->P[h->cur], h->Np, h->K, h->gr, h->red_partial, h->nparts_red, h->red_out, h->stream);
  HIP_TRY(hipGetLastError());
  double o[16];
  HIP_TRY(hipMemcpyAsync(o, h->red_out, sizeof(o), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  out->particles = h->Np;
  out->fluid_cells = (int64_t)o[4];
  for (int a = 0; a < 4; ++a) out->phase_sum[a] = a < h->K ? o[a] : 0.0;
  out->mean_ppc = out->fluid_cells > 0 ? (double)h->Np / (double)out->fluid_cells : 0.0;
  for (int a = 0; a < 3; ++a) {
    out->pos_min[a] = -o[5 + a];
    out->pos_max[a] = o[8 + a];
  }
  return AQ_OK;
}

aq_status aq_flip_solve_pressure(aq_flip* h, double tol, int32_t max_iters, int32_t check_every,
                                 int32_t* iters_out, double* rel_out) {
  if (!h || max_iters < 0 || check_every < 1) return fail(AQ_EINVAL, "bad argument");
  int it = 0;
  double rel = 0;
  AQ_TRY(aq_flip_residual_fp64(h, &rel));
  while (!(rel <= tol) && it < max_iters) {
    const int k = std::min(check_every, max_iters - it);
    AQ_TRY(sweeps_then_residual(h, k, &rel));
    it += k;
  }
  if (iters_out) *iters_out = it;
  if (rel_out) *rel_out = rel;
  if (!(rel <= tol)) return fail(AQ_ESTALL, "pressure solve reached its iteration cap");
  return AQ_OK;
}

// ---- snapshots ---------------------------------------------------------------
extern "C++" {
namespace {

constexpr char kSnapMagic[8] = {'A', 'Q', 'F', 'L', 'I', 'P', 'S', 'N'};

// Every device array that carries state between calls, tagged in a fixed order.
// Ping-pong pairs contribute their current buffer only.
struct SnapArray {
  uint32_t tag;
  void* ptr;
  size_t bytes;
};
std::vector<SnapArray> snapshot_arrays(aq_flip* h) {
  const Geom& G = h->G;
  const size_t nc = (size_t)G.ncell, np = (size_t)h->Np, nh = (size_t)G.ncells_hash;
  Grid& gr = h->gr;
  std::vector<SnapArray> a;
  const Particles& P = h->P[h->cur];
  a.push_back({1, P.a, np * sizeof(float4)});
  a.push_back({2, P.b, np * sizeof(float4)});
  if (h->K > 2) a.push_back({3, P.c, np * sizeof(float4)});
  a.push_back({10, gr.content, nc});
  for (int c = 0; c < 3; ++c) a.push_back({11u + c, gr.vel[c], nc * 4});
  for (int c = 0; c < 3; ++c) a.push_back({14u + c, gr.nvel[h->nv][c], nc * 4});
  a.push_back({17, gr.hasvel[h->hv], nc});
  a.push_back({18, gr.p[h->pb], nc * 4});
  a.push_back({19, gr.div, nc * 4});
  a.push_back({20, gr.rhs, nc * 4});
  a.push_back({21, gr.density, nc * 4});
  for (int c = 0; c < h->K; ++c) a.push_back({30u + c, gr.vf[c], nc * 4});
  for (int c = 0; c < h->K; ++c) a.push_back({40u + c, gr.nvf[h->nf][c], nc * 4});
  a.push_back({50, gr.active, ((nh + 255) / 256 + 1) * sizeof(uint32_t)});
  a.push_back({60, h->B.offsets, (nh + 1) * sizeof(int)});
  a.push_back({61, h->B.hash, np * sizeof(int)});
  a.push_back({62, h->B.perm, np * sizeof(int)});
  return a;
}

// Configuration fields in a fixed order (independent of the struct's padding).
template <typename IO>
void config_fields(IO&& io, aq_flip_config& c) {
  for (int a = 0; a < 3; ++a) io(c.size[a]);
  io(c.dt), io(c.flip_coeff), io(c.particle_jitter), io(c.grid_physical_size);
  io(c.pressure_iters), io(c.diffusion_iters), io(c.phase_count);
  for (int a = 0; a < 3; ++a) io(c.gravity[a]);
  for (int a = 0; a < 4; ++a) io(c.phase_densities[a]);
  io(c.diffusion_coeff), io(c.pressure_init), io(c.ref_boundary_coverage), io(c.seed), io(c.n_volumes);
  for (int i = 0; i < AQ_MAX_VOLUMES; ++i) {
    aq_flip_volume& v = c.volumes[i];
    io(v.shape);
    for (int a = 0; a < 3; ++a) io(v.box_min[a]);
    for (int a = 0; a < 3; ++a) io(v.box_max[a]);
    for (int a = 0; a < 3; ++a) io(v.center[a]);
    io(v.radius), io(v.phase);
  }
}

}  // namespace
}  // extern "C++"

aq_status aq_flip_save(aq_flip* h, const char* path) {
  if (!h || !path) return fail(AQ_EINVAL, "null argument");
  HIP_TRY(hipStreamSynchronize(h->stream));
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(AQ_EINVAL, std::string("cannot open ") + path + " for writing");
  bool ok = true;
  auto put = [&](const void* p, size_t n) { ok = ok && std::fwrite(p, 1, n, f) == n; };
  auto io = [&](auto& v) { put(&v, sizeof(v)); };
  put(kSnapMagic, 8);
  uint32_t version = AQ_SNAPSHOT_VERSION;
  io(version);
  aq_flip_config cfg = h->cfg;
  config_fields(io, cfg);
  int64_t np = h->Np;
  int32_t K = h->K, nsec = 0;
  float t = h->physicalTime;
  io(np), io(K), io(t);
  const std::vector<SnapArray> arrays = snapshot_arrays(h);
  nsec = (int32_t)arrays.size();
  io(nsec);
  std::vector<char> buf;
  for (const SnapArray& a : arrays) {
    uint32_t tag = a.tag;
    uint64_t bytes = a.bytes;
    io(tag), io(bytes);
    buf.resize(a.bytes);
    if (hipMemcpy(buf.data(), a.ptr, a.bytes, hipMemcpyDeviceToHost) != hipSuccess) ok = false;
    put(buf.data(), a.bytes);
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(AQ_EHIP, std::string("writing the snapshot ") + path + " failed");
  return AQ_OK;
}

aq_status aq_flip_load(const char* path, aq_flip** out) {
  if (!path || !out) return fail(AQ_EINVAL, "null argument");
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(AQ_EINVAL, std::string("cannot open ") + path);
  bool ok = true;
  auto get = [&](void* p, size_t n) { ok = ok && std::fread(p, 1, n, f) == n; };
  auto io = [&](auto& v) { get(&v, sizeof(v)); };
  auto bail = [&](aq_status st, const std::string& msg) {
    std::fclose(f);
    return fail(st, msg);
  };
  char magic[8];
  get(magic, 8);
  uint32_t version = 0;
  io(version);
  if (!ok || std::memcmp(magic, kSnapMagic, 8) != 0) return bail(AQ_EINVAL, "not a libaqflip snapshot");
  if (version != AQ_SNAPSHOT_VERSION)
    return bail(AQ_EINVAL, "snapshot format version " + std::to_string(version) + " is not supported");
  aq_flip_config cfg{};
  config_fields(io, cfg);
  int64_t np = 0;
  int32_t K = 0, nsec = 0;
  float t = 0.f;
  io(np), io(K), io(t), io(nsec);
  if (!ok) return bail(AQ_EINVAL, "truncated snapshot header");
  if (validate(&cfg) != AQ_OK) return bail(AQ_EINVAL, std::string("snapshot configuration: ") + aq_last_error());
  if (np <= 0 || np > (1ll << 31) - 1 || K < 1 || K > 4) return bail(AQ_EINVAL, "corrupt snapshot state");
  aq_flip* h = new aq_flip;
  read_env(h);
  aq_status st = setup(h, &cfg, np);
  if (st != AQ_OK) {
    delete h;
    return bail(st, aq_last_error());
  }
  h->K = K;
  h->physicalTime = t;
  const std::vector<SnapArray> arrays = snapshot_arrays(h);
  if ((size_t)nsec != arrays.size()) {
    delete h;
    return bail(AQ_EINVAL, "snapshot sections do not match its configuration");
  }
  std::vector<char> buf;
  for (const SnapArray& a : arrays) {
    uint32_t tag = 0;
    uint64_t bytes = 0;
    io(tag), io(bytes);
    if (!ok || tag != a.tag || bytes != a.bytes) {
      delete h;
      return bail(AQ_EINVAL, "corrupt snapshot section " + std::to_string(a.tag));
    }
    buf.resize(a.bytes);
    get(buf.data(), a.bytes);
    if (!ok || hipMemcpy(a.ptr, buf.data(), a.bytes, hipMemcpyHostToDevice) != hipSuccess) {
      delete h;
      return bail(AQ_EINVAL, "truncated snapshot section " + std::to_string(a.tag));
    }
  }
  std::fclose(f);
  // Pressure buffers: only the current one was saved, so neither is known to hold 0
  // on every AIR cell; the next sweeps write them fully (same values either way).
  h->air_clean[0] = h->air_clean[1] = false;
  launch_active_list(h->G, h->gr, h->stream);  // derived from the restored flags
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    delete h;
    return bail(AQ_EHIP, "snapshot: rebuilding the active list failed");
  }
  *out = h;
  return AQ_OK;
}

aq_status aq_flip_solve_pressure_mgpcg(aq_flip* h, double tol, int32_t max_iters, int32_t check_every,
                                       int32_t from_current, int32_t* iters_out, double* rel_out) {
  if (!h || max_iters < 0 || check_every < 1 || !(tol > 0)) return fail(AQ_EINVAL, "bad argument");
  int it = 0;
  double cg = 0, rel = 0;
  AQ_TRY(mgpcg_solve(h, tol, max_iters, check_every, from_current != 0, &it, &cg));
  AQ_TRY(residual_now(h, &rel));
  if (iters_out) *iters_out = it;
  if (rel_out) *rel_out = rel;
  if (!(cg <= tol)) return fail(AQ_ESTALL, "MG-PCG reached its iteration cap");
  return AQ_OK;
}

aq_status aq_flip_set_timing(aq_flip* h, int32_t enable) {
  if (!h) return fail(AQ_EINVAL, "null handle");
  h->timing = enable != 0;
  return AQ_OK;
}
aq_status aq_flip_get_timing(aq_flip* h, double* stage_ms, int64_t* stage_calls) {
  if (!h) return fail(AQ_EINVAL, "null handle");
  for (int i = 0; i < AQ_STAGE_COUNT; ++i) {
    if (stage_ms) stage_ms[i] = h->stage_ms[i];
    if (stage_calls) stage_calls[i] = h->stage_calls[i];
  }
  return AQ_OK;
}
aq_status aq_flip_paired_diffusion_computed(aq_flip* h, int64_t* count) {
  if (!h || !count) return fail(AQ_EINVAL, "null argument");
  uint32_t v = 0;
  HIP_TRY(hipStreamSynchronize(h->stream));
  HIP_TRY(hipMemcpy(&v, h->gr.dchange + 3, sizeof(v), hipMemcpyDeviceToHost));
  *count = (int64_t)v;
  return AQ_OK;
}

aq_status aq_flip_reset_timing(aq_flip* h) {
  if (!h) return fail(AQ_EINVAL, "null handle");
  for (int i = 0; i < AQ_STAGE_COUNT; ++i) {
    h->stage_ms[i] = 0;
    h->stage_calls[i] = 0;
  }
  return AQ_OK;
}

double aq_flip_physical_time(const aq_flip* h) { return h ? (double)h->physicalTime : 0.0; }

aq_status aq_flip_get_center(const aq_flip* h, float* center) {
  if (!h || !center) return fail(AQ_EINVAL, "null argument");
  // Fluid::getCenter returns gridPhysicalSize/2 on every axis (Fluid_3D_FLIP.cu:728-730).
  for (int a = 0; a < 3; ++a) center[a] = h->cfg.grid_physical_size / 2;
  return AQ_OK;
}
float aq_flip_container_size(const aq_flip* h) { return h ? h->cfg.grid_physical_size : 0.f; }

const char* aq_last_error(void) { return g_err.c_str(); }

const char* aq_flip_version(void) { return "aqflip 0.1.0 (gfx950)"; }

}  // extern "C"
