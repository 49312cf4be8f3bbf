// mgpcg.h -- multigrid-preconditioned CG for the pressure system (mgpcg.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "aqflip_internal.h"

namespace aqf {

// One grid level: cells in hash order (x-major, z fastest), fluid[n] = 1 on unknowns.
struct MgLevel {
  int X, Y, Z;
  long long n;
  uint8_t* fluid;
};

// The V-cycle's levels and FP32 work vectors (f: right-hand side, u: correction,
// t: ping-pong / residual scratch), level 0 being the FLIP grid.
struct MgHierarchy {
  std::vector<MgLevel> levels;
  std::vector<float*> f, u, t;
  int nu = 2;             // damped Jacobi sweeps before and after the coarse correction (even)
  int coarse_sweeps = 40; // sweeps on the coarsest level
};

void mg_gather(const Geom& G, const uint8_t* content, const float* rhs, const float* p, bool use_p,
               const MgLevel& L0, double* b, double* x, hipStream_t s);
void mg_scatter(const Geom& G, const MgLevel& L0, const double* x, float* p, hipStream_t s);
void mg_coarsen_flags(const MgLevel& F, const MgLevel& C, hipStream_t s);
void cg_residual(const MgLevel& L, const double* b, const double* x, double* r, hipStream_t s);
void cg_apply(const MgLevel& L, const double* d, double* q, hipStream_t s);
// out[slot] = a.b, out[slot+1] = c.d (c may be null), deterministic order
void cg_dot2(long long n, const double* a, const double* b, const double* c, const double* d, double* part,
             int nparts, double* out, int slot, hipStream_t s);
// with S[0] = r.z, S[2] = d.q: x += alpha d, r -= alpha q (alpha = S[0]/S[2])
void cg_update_xr(long long n, const double* S, const double* d, const double* q, double* x, double* r,
                  hipStream_t s);
// d = z + beta d with beta = S[4]/S[0]
void cg_update_d(long long n, const double* S, const double* z, double* d, hipStream_t s);
// S[0] = S[4], S[1] = S[5]
void cg_shift(double* S, hipStream_t s);
void cg_copy(long long n, const double* a, double* b, hipStream_t s);
// z = M^-1 r (one V-cycle)
void mg_vcycle(const MgHierarchy& H, const double* r, double* z, hipStream_t s);

}  // namespace aqf
