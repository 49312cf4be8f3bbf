// mgpcg.hip -- multigrid-preconditioned conjugate gradients for the pressure system.
//
// The accuracy reference and second fallback beside Jacobi. It solves the same
// system as jacobiImpl's fixed point (Fluid/Fluid_3D_kernels.cu:460-505) -- the
// operator of the FP64 residual, (A p)_i = (6 - walls_i) p_i - sum of FLUID
// neighbours, AIR cells Dirichlet 0 -- which is the matrix the reference's legacy
// sparse path assembles (Fluid/Legacy/PressureEquation.cuh:216-328) and solves with
// AMGX-preconditioned CG (Fluid/Legacy/SPD_Solver.h:102-160). Here it is matrix-free:
// CG in FP64 on dense vectors in hash order, preconditioned by one geometric V-cycle
// in FP32 (levels halve every axis; a coarse cell is FLUID when any child is; damped
// Jacobi, omega 2/3, nu sweeps before and after; cell-centred trilinear prolongation
// and its exact transpose as restriction, scaled by 1/2 for the coarser operator; a
// fixed number of sweeps on the coarsest level). Every piece is a fixed linear map
// and pre/post smoothing match, so the preconditioner is symmetric positive definite
// as CG requires. Scalars (alpha, beta) stay on the device; the host reads the
// residual norm only every `check_every` iterations.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "aqflip_internal.h"
#include "mgpcg.h"

namespace aqf {

constexpr int kMgBlock = 256;
static inline int mg_blocks(long long n) { return (int)((n + kMgBlock - 1) / kMgBlock); }

__device__ __forceinline__ void mg_cell(const MgLevel& L, long long n, int& x, int& y, int& z) {
  z = (int)(n % L.Z);
  const long long t = n / L.Z;
  y = (int)(t % L.Y);
  x = (int)(t / L.Y);
}

// (A u)_n on a level (unit spacing): walls are clamp-to-edge (Neumann), non-FLUID
// neighbours are Dirichlet 0.
template <typename T>
__device__ __forceinline__ T apply_A(const MgLevel& L, const T* __restrict__ u, long long n, int x, int y,
                                     int z) {
  const T uc = u[n];
  T a = (T)6 * uc;
  const long long sx = (long long)L.Y * L.Z, sy = L.Z;
  if (x + 1 < L.X) { if (L.fluid[n + sx]) a -= u[n + sx]; } else a -= uc;
  if (x > 0)       { if (L.fluid[n - sx]) a -= u[n - sx]; } else a -= uc;
  if (y + 1 < L.Y) { if (L.fluid[n + sy]) a -= u[n + sy]; } else a -= uc;
  if (y > 0)       { if (L.fluid[n - sy]) a -= u[n - sy]; } else a -= uc;
  if (z + 1 < L.Z) { if (L.fluid[n + 1]) a -= u[n + 1]; } else a -= uc;
  if (z > 0)       { if (L.fluid[n - 1]) a -= u[n - 1]; } else a -= uc;
  return a;
}
__device__ __forceinline__ int walls_of(const MgLevel& L, int x, int y, int z) {
  return (x + 1 >= L.X) + (x <= 0) + (y + 1 >= L.Y) + (y <= 0) + (z + 1 >= L.Z) + (z <= 0);
}

// ---- level 0 <-> the FLIP grid
__global__ void k_mg_gather(Geom G, const uint8_t* __restrict__ content, const float* __restrict__ rhs,
                            const float* __restrict__ p, int use_p, MgLevel L0, double* __restrict__ b,
                            double* __restrict__ x) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= L0.n) return;
  int i, j, k;
  mg_cell(L0, n, i, j, k);
  const long long g = G.g(i, j, k);
  const bool fl = content[g] == FLUID;
  L0.fluid[n] = fl ? 1 : 0;
  b[n] = fl ? (double)rhs[g] : 0.0;
  x[n] = (fl && use_p) ? (double)p[g] : 0.0;
}
__global__ void k_mg_scatter(Geom G, MgLevel L0, const double* __restrict__ x, float* __restrict__ p) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= L0.n) return;
  int i, j, k;
  mg_cell(L0, n, i, j, k);
  p[G.g(i, j, k)] = L0.fluid[n] ? (float)x[n] : 0.f;
}

// ---- coarse flags: FLUID when any child is
__global__ void k_mg_coarsen_flags(MgLevel F, MgLevel C) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= C.n) return;
  int x, y, z;
  mg_cell(C, n, x, y, z);
  uint8_t any = 0;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        const int fx = 2 * x + a, fy = 2 * y + b, fz = 2 * z + c;
        if (fx < F.X && fy < F.Y && fz < F.Z) any |= F.fluid[((long long)fx * F.Y + fy) * F.Z + fz];
      }
  C.fluid[n] = any;
}

// ---- CG vector kernels (FP64)
__global__ void k_cg_residual(MgLevel L, const double* __restrict__ b, const double* __restrict__ x,
                              double* __restrict__ r) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= L.n) return;
  int i, j, k;
  mg_cell(L, n, i, j, k);
  r[n] = L.fluid[n] ? b[n] - apply_A<double>(L, x, n, i, j, k) : 0.0;
}
__global__ void k_cg_apply(MgLevel L, const double* __restrict__ d, double* __restrict__ q) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= L.n) return;
  int i, j, k;
  mg_cell(L, n, i, j, k);
  q[n] = L.fluid[n] ? apply_A<double>(L, d, n, i, j, k) : 0.0;
}
// Deterministic dot products: per-block partial sums of up to two products, folded
// in block order by one block.
__global__ void k_cg_dot2(long long n, const double* __restrict__ a, const double* __restrict__ b,
                          const double* __restrict__ c, const double* __restrict__ d,
                          double* __restrict__ part) {
  __shared__ double s1[kMgBlock], s2[kMgBlock];
  double v1 = 0, v2 = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    v1 += a[i] * b[i];
    if (c) v2 += c[i] * d[i];
  }
  s1[threadIdx.x] = v1;
  s2[threadIdx.x] = v2;
  __syncthreads();
  for (int o = kMgBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = s2[0];
  }
}
// out[slot], out[slot+1] = folded sums
__global__ void k_cg_fold(const double* __restrict__ part, int nparts, double* __restrict__ out, int slot) {
  if (threadIdx.x != 0) return;
  double a = 0, b = 0;
  for (int i = 0; i < nparts; ++i) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  out[slot] = a;
  out[slot + 1] = b;
}
// Scalars in S: [0] rz, [1] rr, [2] dq, [3] (unused), [4] rz_new, [5] rr_new.
// x += alpha d, r -= alpha q with alpha = rz / dq.
__global__ void k_cg_update_xr(long long n, const double* __restrict__ S, const double* __restrict__ d,
                               const double* __restrict__ q, double* __restrict__ x, double* __restrict__ r) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double alpha = S[2] != 0.0 ? S[0] / S[2] : 0.0;
  x[i] += alpha * d[i];
  r[i] -= alpha * q[i];
}
// d = z + beta d with beta = rz_new / rz.
__global__ void k_cg_update_d(long long n, const double* __restrict__ S, const double* __restrict__ z,
                              double* __restrict__ d) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double beta = S[0] != 0.0 ? S[4] / S[0] : 0.0;
  d[i] = z[i] + beta * d[i];
}
__global__ void k_cg_shift(double* __restrict__ S) {
  if (threadIdx.x == 0) {
    S[0] = S[4];
    S[1] = S[5];
  }
}
__global__ void k_copy_d(long long n, const double* __restrict__ a, double* __restrict__ b) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// ---- V-cycle kernels (FP32)
__global__ void k_mg_to_f32(long long n, const double* __restrict__ r, float* __restrict__ f) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = (float)r[i];
}
__global__ void k_mg_to_f64(long long n, const float* __restrict__ u, double* __restrict__ z) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = (double)u[i];
}
// Damped Jacobi: out = in + omega (f - A in) / diag on FLUID cells; `zero_in` treats
// `in` as all zeros (the first sweep of a level).
__global__ void k_mg_smooth(MgLevel L, const float* __restrict__ f, const float* __restrict__ in,
                            float* __restrict__ out, int zero_in) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= L.n) return;
  if (!L.fluid[n]) {
    out[n] = 0.f;
    return;
  }
  int x, y, z;
  mg_cell(L, n, x, y, z);
  const float diag = (float)(6 - walls_of(L, x, y, z));
  const float omega = 2.f / 3.f;
  if (zero_in) {
    out[n] = omega * f[n] / diag;
    return;
  }
  const float r = f[n] - apply_A<float>(L, in, n, x, y, z);
  out[n] = in[n] + omega * r / diag;
}
__global__ void k_mg_residual(MgLevel L, const float* __restrict__ f, const float* __restrict__ u,
                              float* __restrict__ res) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= L.n) return;
  int x, y, z;
  mg_cell(L, n, x, y, z);
  res[n] = L.fluid[n] ? f[n] - apply_A<float>(L, u, n, x, y, z) : 0.f;
}
// Cell-centred trilinear prolongation, per axis: fine x takes 3/4 of parent p = x/2
// and 1/4 of q = p + (x odd ? 1 : -1) clamped to the coarse range.
__device__ __forceinline__ void prolong_refs(int x, int nc, int& p, int& q) {
  p = min(x >> 1, nc - 1);
  q = min(max(p + ((x & 1) ? 1 : -1), 0), nc - 1);
}
__device__ __forceinline__ float axis_weight(int x, int a, int nc) {
  int p, q;
  prolong_refs(x, nc, p, q);
  return (p == a ? 0.75f : 0.f) + (q == a ? 0.25f : 0.f);
}
// fine u += P e on FLUID fine cells
__global__ void k_mg_prolong_add(MgLevel F, MgLevel C, const float* __restrict__ e, float* __restrict__ u) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= F.n) return;
  if (!F.fluid[n]) return;
  int x, y, z;
  mg_cell(F, n, x, y, z);
  int px, qx, py, qy, pz, qz;
  prolong_refs(x, C.X, px, qx);
  prolong_refs(y, C.Y, py, qy);
  prolong_refs(z, C.Z, pz, qz);
  const int ix[2] = {px, qx}, iy[2] = {py, qy}, iz[2] = {pz, qz};
  const float w[2] = {0.75f, 0.25f};
  float s = 0.f;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        const long long m = ((long long)ix[a] * C.Y + iy[b]) * C.Z + iz[c];
        if (C.fluid[m]) s += w[a] * w[b] * w[c] * e[m];
      }
  u[n] += s;
}
// coarse f = 1/2 * P^T (fine residual), the exact transpose of k_mg_prolong_add
// (same masks), gathered: coarse a collects the fine cells 2a-1 .. 2a+2 per axis.
__global__ void k_mg_restrict(MgLevel F, MgLevel C, const float* __restrict__ res, float* __restrict__ f) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= C.n) return;
  if (!C.fluid[n]) {
    f[n] = 0.f;
    return;
  }
  int x, y, z;
  mg_cell(C, n, x, y, z);
  float s = 0.f;
  for (int fx = max(2 * x - 1, 0); fx <= min(2 * x + 2, F.X - 1); ++fx) {
    const float wx = axis_weight(fx, x, C.X);
    if (wx == 0.f) continue;
    for (int fy = max(2 * y - 1, 0); fy <= min(2 * y + 2, F.Y - 1); ++fy) {
      const float wy = axis_weight(fy, y, C.Y);
      if (wy == 0.f) continue;
      for (int fz = max(2 * z - 1, 0); fz <= min(2 * z + 2, F.Z - 1); ++fz) {
        const float wz = axis_weight(fz, z, C.Z);
        if (wz == 0.f) continue;
        const long long m = ((long long)fx * F.Y + fy) * F.Z + fz;
        if (F.fluid[m]) s += wx * wy * wz * res[m];
      }
    }
  }
  f[n] = 0.5f * s;
}

// ---- launchers
void mg_gather(const Geom& G, const uint8_t* content, const float* rhs, const float* p, bool use_p,
               const MgLevel& L0, double* b, double* x, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_gather, dim3(mg_blocks(L0.n)), dim3(kMgBlock), 0, s, G, content, rhs, p, use_p ? 1 : 0,
                     L0, b, x);
}
void mg_scatter(const Geom& G, const MgLevel& L0, const double* x, float* p, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_scatter, dim3(mg_blocks(L0.n)), dim3(kMgBlock), 0, s, G, L0, x, p);
}
void mg_coarsen_flags(const MgLevel& F, const MgLevel& C, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_coarsen_flags, dim3(mg_blocks(C.n)), dim3(kMgBlock), 0, s, F, C);
}
void cg_residual(const MgLevel& L, const double* b, const double* x, double* r, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_residual, dim3(mg_blocks(L.n)), dim3(kMgBlock), 0, s, L, b, x, r);
}
void cg_apply(const MgLevel& L, const double* d, double* q, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_apply, dim3(mg_blocks(L.n)), dim3(kMgBlock), 0, s, L, d, q);
}
void cg_dot2(long long n, const double* a, const double* b, const double* c, const double* d, double* part,
             int nparts, double* out, int slot, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_dot2, dim3(nparts), dim3(kMgBlock), 0, s, n, a, b, c, d, part);
  hipLaunchKernelGGL(k_cg_fold, dim3(1), dim3(64), 0, s, part, nparts, out, slot);
}
void cg_update_xr(long long n, const double* S, const double* d, const double* q, double* x, double* r,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_cg_update_xr, dim3(mg_blocks(n)), dim3(kMgBlock), 0, s, n, S, d, q, x, r);
}
void cg_update_d(long long n, const double* S, const double* z, double* d, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_update_d, dim3(mg_blocks(n)), dim3(kMgBlock), 0, s, n, S, z, d);
}
void cg_shift(double* S, hipStream_t s) { hipLaunchKernelGGL(k_cg_shift, dim3(1), dim3(64), 0, s, S); }
void cg_copy(long long n, const double* a, double* b, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_d, dim3(mg_blocks(n)), dim3(kMgBlock), 0, s, n, a, b);
}

// z = V-cycle(r) on the level hierarchy (FP32 inside). Every smoothing sweep reads
// one of u[l], t[l] and writes the other; the sweep counts are arranged so that
// each phase ends in u[l].
void mg_vcycle(const MgHierarchy& H, const double* r, double* z, hipStream_t s) {
  const int nl = (int)H.levels.size();
  const MgLevel& L0 = H.levels[0];
  hipLaunchKernelGGL(k_mg_to_f32, dim3(mg_blocks(L0.n)), dim3(kMgBlock), 0, s, L0.n, r, H.f[0]);
  for (int l = 0; l < nl; ++l) {
    const MgLevel& L = H.levels[l];
    const int sweeps = (l == nl - 1) ? H.coarse_sweeps : H.nu;
    for (int k = 0; k < sweeps; ++k) {  // from zero; the last sweep writes u[l]
      const bool to_u = (sweeps - 1 - k) % 2 == 0;
      float* out = to_u ? H.u[l] : H.t[l];
      const float* in = to_u ? H.t[l] : H.u[l];
      hipLaunchKernelGGL(k_mg_smooth, dim3(mg_blocks(L.n)), dim3(kMgBlock), 0, s, L, H.f[l], in, out,
                         k == 0 ? 1 : 0);
    }
    if (l == nl - 1) break;
    const MgLevel& C = H.levels[l + 1];
    hipLaunchKernelGGL(k_mg_residual, dim3(mg_blocks(L.n)), dim3(kMgBlock), 0, s, L, H.f[l], H.u[l], H.t[l]);
    hipLaunchKernelGGL(k_mg_restrict, dim3(mg_blocks(C.n)), dim3(kMgBlock), 0, s, L, C, H.t[l], H.f[l + 1]);
  }
  for (int l = nl - 2; l >= 0; --l) {
    const MgLevel& L = H.levels[l];
    const MgLevel& C = H.levels[l + 1];
    hipLaunchKernelGGL(k_mg_prolong_add, dim3(mg_blocks(L.n)), dim3(kMgBlock), 0, s, L, C, H.u[l + 1], H.u[l]);
    for (int k = 0; k < H.nu; ++k) {  // nu is even: u -> t -> u
      const bool even = (k % 2 == 0);
      hipLaunchKernelGGL(k_mg_smooth, dim3(mg_blocks(L.n)), dim3(kMgBlock), 0, s, L, H.f[l],
                         even ? H.u[l] : H.t[l], even ? H.t[l] : H.u[l], 0);
    }
  }
  hipLaunchKernelGGL(k_mg_to_f64, dim3(mg_blocks(L0.n)), dim3(kMgBlock), 0, s, L0.n, H.u[0], z);
}

}  // namespace aqf
