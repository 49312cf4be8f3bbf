// unet_kernels.hip -- BF16 MFMA kernels of the pressure warm-start U-Net (gfx950).
//
// Tensors are channels-last (B, D, H, W, C) in bf16 with a channel stride and
// offset, so a decoder's concatenated input is one buffer that the skip
// connection and the upsampled path write into side by side. All three 3x3x3
// convolution products -- forward, input gradient and weight gradient -- run on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
//
// MFMA 16x16x32 lane layout (wave64): lane l supplies row (l % 16) of A and
// column (l % 16) of B, each with the 8 consecutive k values 8*(l/16) .. +7; it
// receives C rows 4*(l/16) .. +3 of column (l % 16).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstdlib>

#include "unet_internal.h"

namespace aqu {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kXCD = 8;

__device__ __forceinline__ float bf2f(__bf16 v) { return (float)v; }
__device__ __forceinline__ __bf16 f2bf(float v) { return (__bf16)v; }  // round to nearest even

__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, bid = blockIdx.x;
  const int per = nb / kXCD, rem = nb % kXCD;
  const int xcd = bid % kXCD, k = bid / kXCD;
  return xcd * per + min(xcd, rem) + k;
}

struct VoxelPos {
  int b, d, h, w;
};
__device__ __forceinline__ VoxelPos vox_of(const Vol& v, long long id) {
  VoxelPos p;
  p.w = (int)(id % v.W);
  long long t = id / v.W;
  p.h = (int)(t % v.H);
  t /= v.H;
  p.d = (int)(t % v.D);
  p.b = (int)(t / v.D);
  return p;
}
__device__ __forceinline__ long long vox_id(const Vol& v, int b, int d, int h, int w) {
  return (((long long)b * v.D + d) * v.H + h) * v.W + w;
}

// ------------------------------------------------------------------ conv 3x3x3, implicit GEMM
// Y[v, n] = sum_{tap, c} X[v + off(tap), c] * Wk[n, tap*CIN + c] (+ bias, ReLU),
// zero padding outside the volume. M = voxels, N = COUT, K = 27*CIN in slices of
// 32. A wave owns kConvMF fragments of 16 consecutive voxels, so every weight
// fragment it fetches feeds kConvMF MFMAs and the fragments' input gathers are in
// flight together. When the layer's weights fit (<= 64 KiB) the block stages them
// in LDS once, rows padded by 16 bytes so the 16 rows a wave reads fall in
// distinct banks. With flipped, transposed weights the same kernel computes the
// input gradient. OUT_F32 writes FP32 (optionally accumulating), else bf16.
// AQUNET_NO_TILE=1 switches the LDS-tiled forward convolution off (measurements and
// the tiled-against-untiled test); read at every launch.
static bool conv3_tile_enabled() {
  const char* e = std::getenv("AQUNET_NO_TILE");
  return !(e && e[0] == '1');
}

constexpr int kConvMF = 4;
constexpr int kConvVox = 4 * kConvMF * 16;  // voxels per 256-thread block

// MF fragments of 16 voxels per wave; the block computes NTW 16-channel output
// tiles, tile group blockIdx.y (NTW = COUT/16: all of them). Small grids use MF = 1,
// NTW = 1 so that they still spread over the CUs; each output value is the same
// MFMA chain either way.
template <int CIN, int COUT, bool RELU, bool OUT_F32, bool ACCUM, int MF = kConvMF, int NTW = COUT / 16>
__global__ __launch_bounds__(256) void k_conv3(Vol v, const __bf16* __restrict__ X, int xs, int xo,
                                               const __bf16* __restrict__ Wk,
                                               const float* __restrict__ bias, void* __restrict__ Yv,
                                               int ys, int yo) {
  constexpr int K = 27 * CIN;
  constexpr int NT = NTW;
  constexpr int RS = K + 8;  // LDS row stride in bf16
  constexpr bool WLDS = NT * 16 * RS * 2 <= 64 * 1024;
  __shared__ __align__(16) __bf16 sW[WLDS ? NT * 16 * RS : 8];
  const int n0 = (int)blockIdx.y * NT * 16;  // first output channel of this block
  Wk += (long long)n0 * K;
  if (bias) bias += n0;
  if (WLDS) {
    for (int e = threadIdx.x; e < NT * 16 * (K / 8); e += 256) {
      const int n = e / (K / 8), j = e - n * (K / 8);
      *(bf16x8*)&sW[n * RS + j * 8] = *(const bf16x8*)(Wk + (long long)n * K + j * 8);
    }
    __syncthreads();
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long V = (long long)v.B * v.D * v.H * v.W;
  const long long wbase = ((long long)xcd_block() * 4 + wave) * (MF * 16);
  if (wbase >= V) return;
  const int r = lane & 15, q = lane >> 4;
  VoxelPos p[MF];
  bool vok[MF];
#pragma unroll
  for (int f = 0; f < MF; ++f) {
    const long long vid = wbase + f * 16 + r;
    vok[f] = vid < V;
    p[f] = vox_of(v, vok[f] ? vid : 0);
  }
  f32x4 acc[MF][NT];
#pragma unroll
  for (int f = 0; f < MF; ++f)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[f][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16x8 zero8 = {};
  for (int k0 = 0; k0 < K; k0 += 32) {
    const int k = k0 + q * 8;
    const bool kok = k < K;
    const int tap = k / CIN, c = k - tap * CIN;
    const int dd = tap / 9 - 1, dh = (tap / 3) % 3 - 1, dw = tap % 3 - 1;
    bf16x8 a[MF];
#pragma unroll
    for (int f = 0; f < MF; ++f) {
      a[f] = zero8;
      const int zd = p[f].d + dd, zh = p[f].h + dh, zw = p[f].w + dw;
      if (kok && vok[f] && zd >= 0 && zd < v.D && zh >= 0 && zh < v.H && zw >= 0 && zw < v.W)
        a[f] = *(const bf16x8*)(X + vox_id(v, p[f].b, zd, zh, zw) * xs + xo + c);
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      bf16x8 bw = zero8;
      if (kok) {
        if (WLDS)
          bw = *(const bf16x8*)&sW[(t * 16 + r) * RS + k];
        else
          bw = *(const bf16x8*)(Wk + (long long)(t * 16 + r) * K + k);
      }
#pragma unroll
      for (int f = 0; f < MF; ++f)
        acc[f][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[f], bw, acc[f][t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int f = 0; f < MF; ++f)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int n = t * 16 + r;  // within this block's tiles
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const long long o = wbase + f * 16 + q * 4 + i;
        if (o >= V) continue;
        float val = acc[f][t][i];
        if (bias) val += bias[n];
        if (RELU) val = fmaxf(val, 0.f);
        if (OUT_F32) {
          float* Y = (float*)Yv + o * ys + yo + n0 + n;
          *Y = ACCUM ? *Y + val : val;
        } else {
          ((__bf16*)Yv)[o * ys + yo + n0 + n] = f2bf(val);
        }
      }
    }
}

// ------------------------------------------------------------------ conv 3x3x3, input tile in LDS
// The forward convolution of the layers whose extents divide into tiles: a block owns
// an output tile of TD x 8 x 16 voxels (d, h, w) of one batch item and NTB 16-channel
// output tiles (channel group blockIdx.y). The input tile with its one-voxel halo,
// (TD+2) x 10 x 18 voxels, is staged in LDS CH input channels at a time (zero outside
// the volume), so each input value is read from memory about 2.5 times per layer
// instead of once per tap; the A fragments of all 27 taps then come from LDS. Each of
// the four waves owns MF = 2*TD rows of 16 voxels along w. Voxels sit CH + 8 bf16
// apart in LDS (CH when CH = 8), which puts the 16-byte reads of 8 consecutive lanes in
// distinct banks. The k order is (chunk, tap, channel within the chunk): for CIN = CH
// it is the order of k_conv3.
template <int CIN, int COUT, int CH, int TD, int NTB, bool RELU>
__global__ __launch_bounds__(256) void k_conv3_tile(Vol v, const __bf16* __restrict__ X, int xs, int xo,
                                                    const __bf16* __restrict__ Wk,
                                                    const float* __restrict__ bias, __bf16* __restrict__ Y,
                                                    int ys, int yo) {
  constexpr int TH = 8, TW = 16;
  constexpr int ID = TD + 2, IH = TH + 2, IW = TW + 2;
  constexpr int NV = ID * IH * IW;
  constexpr int VS = CH == 8 ? 8 : CH + 8;  // LDS voxel stride (bf16)
  constexpr int MF = TD * TH / 4;           // 16-voxel rows per wave
  constexpr int KC = 27 * CH;               // k per channel chunk
  constexpr int NSL = (KC + 31) / 32;
  constexpr int K = 27 * CIN;
  static_assert(CIN % CH == 0 && CH % 8 == 0, "channel chunks of 8, 16 or 32");
  __shared__ __align__(16) __bf16 sX[NV * VS];
  const int ntw = v.W / TW, nth = v.H / TH, ntd = v.D / TD;
  int t = xcd_block();
  const int tw = t % ntw;
  t /= ntw;
  const int th = t % nth;
  t /= nth;
  const int td = t % ntd;
  const int b = t / ntd;
  const int d0 = td * TD, h0 = th * TH, w0 = tw * TW;
  const int n0 = (int)blockIdx.y * NTB * 16;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  f32x4 acc[MF][NTB];
#pragma unroll
  for (int f = 0; f < MF; ++f)
#pragma unroll
    for (int nt = 0; nt < NTB; ++nt) acc[f][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16x8 zero8 = {};
  for (int cc = 0; cc < CIN; cc += CH) {
    if (cc) __syncthreads();  // every wave is done with the previous chunk
    for (int e = threadIdx.x; e < NV * (CH / 8); e += 256) {
      const int vx = e / (CH / 8), j = e - vx * (CH / 8);
      const int iw = vx % IW, ih = (vx / IW) % IH, id = vx / (IW * IH);
      const int zd = d0 + id - 1, zh = h0 + ih - 1, zw = w0 + iw - 1;
      bf16x8 val = zero8;
      if (zd >= 0 && zd < v.D && zh >= 0 && zh < v.H && zw >= 0 && zw < v.W)
        val = *(const bf16x8*)(X + vox_id(v, b, zd, zh, zw) * xs + xo + cc + j * 8);
      *(bf16x8*)&sX[vx * VS + j * 8] = val;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NSL; ++s) {
      const int kc = s * 32 + q * 8;
      const bool kok = kc < KC;
      const int tap = kok ? kc / CH : 0, c = kc - tap * CH;
      const int dd = tap / 9, dh = (tap / 3) % 3, dw = tap % 3;
      bf16x8 bw[NTB];
#pragma unroll
      for (int nt = 0; nt < NTB; ++nt)
        bw[nt] = kok ? *(const bf16x8*)(Wk + (long long)(n0 + nt * 16 + r) * K + tap * CIN + cc + c) : zero8;
#pragma unroll
      for (int f = 0; f < MF; ++f) {
        const int row = wave * MF + f;
        const int rd = row / TH, rh = row - rd * TH;
        const bf16x8 a =
            kok ? *(const bf16x8*)&sX[(((rd + dd) * IH + (rh + dh)) * IW + (r + dw)) * VS + c] : zero8;
#pragma unroll
        for (int nt = 0; nt < NTB; ++nt)
          acc[f][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[nt], acc[f][nt], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int f = 0; f < MF; ++f) {
    const int row = wave * MF + f;
    const int rd = row / TH, rh = row - rd * TH;
#pragma unroll
    for (int nt = 0; nt < NTB; ++nt) {
      const int n = n0 + nt * 16 + r;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const long long o = vox_id(v, b, d0 + rd, h0 + rh, w0 + q * 4 + i);
        float val = acc[f][nt][i];
        if (bias) val += bias[n];
        if (RELU) val = fmaxf(val, 0.f);
        Y[o * ys + yo + n] = f2bf(val);
      }
    }
  }
}

// Launches k_conv3_tile when the extents divide into its tiles (W % 16, H % 8 and,
// for TD = 2, D % 2), with ReLU. Two depths TD and one or two channel tiles per
// block, chosen so that the grid still has about 512 or more blocks.
template <int CIN, int COUT, int CH>
static bool conv3_tile_fwd(const Vol& v, const __bf16* X, int xs, int xo, const __bf16* Wk, const float* bias,
                           __bf16* Y, int ys, int yo, hipStream_t s) {
  if (v.W % 16 || v.H % 8) return false;
  const long long tiles1 = (long long)v.B * v.D * (v.H / 8) * (v.W / 16);
  const bool td2 = v.D % 2 == 0 && tiles1 / 2 * (COUT / 16) >= 512;
  const long long tiles = td2 ? tiles1 / 2 : tiles1;
  const bool ntb2 = COUT >= 32 && tiles * (COUT / 32) >= 512;
  const dim3 grid((unsigned)tiles, ntb2 ? COUT / 32 : COUT / 16);
  if (td2 && ntb2)
    hipLaunchKernelGGL((k_conv3_tile<CIN, COUT, CH, 2, 2, true>), grid, dim3(256), 0, s, v, X, xs, xo, Wk, bias, Y, ys, yo);
  else if (td2)
    hipLaunchKernelGGL((k_conv3_tile<CIN, COUT, CH, 2, 1, true>), grid, dim3(256), 0, s, v, X, xs, xo, Wk, bias, Y, ys, yo);
  else if (ntb2)
    hipLaunchKernelGGL((k_conv3_tile<CIN, COUT, CH, 1, 2, true>), grid, dim3(256), 0, s, v, X, xs, xo, Wk, bias, Y, ys, yo);
  else
    hipLaunchKernelGGL((k_conv3_tile<CIN, COUT, CH, 1, 1, true>), grid, dim3(256), 0, s, v, X, xs, xo, Wk, bias, Y, ys, yo);
  return true;
}

template <int CIN, int COUT>
static void conv3_fwd_t(const Vol& v, const __bf16* X, int xs, int xo, const __bf16* Wk,
                        const float* bias, __bf16* Y, int ys, int yo, bool relu, hipStream_t s) {
  constexpr int CH = CIN % 32 == 0 ? 32 : CIN % 16 == 0 ? 16 : 8;
  if (relu && conv3_tile_enabled() && conv3_tile_fwd<CIN, COUT, CH>(v, X, xs, xo, Wk, bias, Y, ys, yo, s)) return;
  const long long V = (long long)v.B * v.D * v.H * v.W;
  if (V >= 1024 * (long long)kConvVox) {  // enough blocks: 4 fragments per wave, all tiles
    const int nb = (int)((V + kConvVox - 1) / kConvVox);
    if (relu)
      hipLaunchKernelGGL((k_conv3<CIN, COUT, true, false, false>), dim3(nb), dim3(256), 0, s, v, X, xs, xo, Wk, bias, (void*)Y, ys, yo);
    else
      hipLaunchKernelGGL((k_conv3<CIN, COUT, false, false, false>), dim3(nb), dim3(256), 0, s, v, X, xs, xo, Wk, bias, (void*)Y, ys, yo);
  } else {  // small grids: one fragment per wave, one 16-channel tile per block
    const dim3 grid((unsigned)((V + 63) / 64), COUT / 16);
    if (relu)
      hipLaunchKernelGGL((k_conv3<CIN, COUT, true, false, false, 1, 1>), grid, dim3(256), 0, s, v, X, xs, xo, Wk, bias, (void*)Y, ys, yo);
    else
      hipLaunchKernelGGL((k_conv3<CIN, COUT, false, false, false, 1, 1>), grid, dim3(256), 0, s, v, X, xs, xo, Wk, bias, (void*)Y, ys, yo);
  }
}
template <int CIN, int COUT>
static void conv3_dgrad_t(const Vol& v, const __bf16* dY, const __bf16* Wt, float* dX, int dxs, int dxo,
                          hipStream_t s) {
  // transposed problem: input dYpre has CIN (= the layer's COUT) channels, output COUT (= its CIN)
  const long long V = (long long)v.B * v.D * v.H * v.W;
  const int nb = (int)((V + kConvVox - 1) / kConvVox);
  hipLaunchKernelGGL((k_conv3<CIN, COUT, false, true, false>), dim3(nb), dim3(256), 0, s, v, dY, CIN, 0, Wt, (const float*)nullptr, (void*)dX, dxs, dxo);
}

void conv3_forward(int cin, int cout, const Vol& v, const __bf16* X, int xs, int xo, const __bf16* Wk,
                   const float* bias, __bf16* Y, int ys, int yo, bool relu, hipStream_t s) {
#define AQU_FWD(CI, CO) \
  if (cin == CI && cout == CO) return conv3_fwd_t<CI, CO>(v, X, xs, xo, Wk, bias, Y, ys, yo, relu, s);
  AQU_FWD(8, 16) AQU_FWD(16, 16) AQU_FWD(16, 32) AQU_FWD(32, 32) AQU_FWD(32, 64) AQU_FWD(64, 64)
  AQU_FWD(96, 32) AQU_FWD(48, 16)
#undef AQU_FWD
}

// Input gradient: dX[v, c] = sum_{tap, n} dYpre[v - off(tap), n] * W[n, tap, c], computed as a
// forward convolution of dYpre (cout channels) with Wt[c, tap', n] = W[n, 26 - tap', c].
void conv3_dgrad(int cin, int cout, const Vol& v, const __bf16* dY, const __bf16* Wt, float* dX,
                 int dxs, int dxo, hipStream_t s) {
#define AQU_DG(CI, CO) \
  if (cin == CI && cout == CO) return conv3_dgrad_t<CO, CI>(v, dY, Wt, dX, dxs, dxo, s);
  AQU_DG(16, 16) AQU_DG(16, 32) AQU_DG(32, 32) AQU_DG(32, 64) AQU_DG(64, 64) AQU_DG(96, 32)
  AQU_DG(48, 16)
#undef AQU_DG
}

// ------------------------------------------------------------------ weight gradient
// dW[n, tap, c] = sum_v dYpre[v, n] * X[v + off(tap), c]: per tap an MFMA GEMM
// (COUT x CIN) with K = voxels. A block owns one tap and a range of W-rows; for
// each row it stages dYpre and the tap-shifted X row into LDS transposed
// (channel-major, voxel contiguous), so a lane's 8 consecutive k (voxels) are one
// aligned 16-byte LDS read. Partial sums per block go to `part`
// [split][tap][COUT][CIN] and are folded in a fixed order by k_fold_wgrad.
constexpr int kWgRow = 64;  // max W handled by the row staging (rows padded to 64)

template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv3_wgrad(Vol v, const __bf16* __restrict__ dY,
                                                     const __bf16* __restrict__ X, int xs, int xo,
                                                     float* __restrict__ part, int nsplit) {
  constexpr int CINP = CIN < 16 ? 16 : CIN;  // N tiles of 16 (first layer has 8 channels)
  constexpr int MT = COUT / 16, NTL = CINP / 16, TILES = MT * NTL;
  constexpr int TPW = (TILES + 3) / 4;  // tiles per wave
  __shared__ __align__(16) __bf16 sDY[COUT][kWgRow];
  __shared__ __align__(16) __bf16 sX[CINP][kWgRow];
  const int split = blockIdx.x, tap = blockIdx.y;
  const int dd = tap / 9 - 1, dh = (tap / 3) % 3 - 1, dw = tap % 3 - 1;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  const long long nrows = (long long)v.B * v.D * v.H;
  f32x4 acc[TPW];
#pragma unroll
  for (int i = 0; i < TPW; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int kch = (v.W + 31) / 32;
  for (long long row = split; row < nrows; row += nsplit) {
    const int h = (int)(row % v.H);
    const long long t = row / v.H;
    const int d = (int)(t % v.D), b = (int)(t / v.D);
    const int sd = d + dd, sh = h + dh;
    const bool rowok = sd >= 0 && sd < v.D && sh >= 0 && sh < v.H;
    __syncthreads();
    for (int e = threadIdx.x; e < COUT * kWgRow; e += 256) {
      const int w = e / COUT, n = e - w * COUT;
      __bf16 val = f2bf(0.f);
      if (w < v.W) val = dY[vox_id(v, b, d, h, w) * COUT + n];
      sDY[n][w] = val;
    }
    for (int e = threadIdx.x; e < CINP * kWgRow; e += 256) {
      const int w = e / CINP, c = e - w * CINP;
      __bf16 val = f2bf(0.f);
      const int sw = w + dw;
      if (rowok && c < CIN && w < v.W && sw >= 0 && sw < v.W)
        val = X[vox_id(v, b, sd, sh, sw) * xs + xo + c];
      sX[c][w] = val;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
      const int tile = wave + 4 * i;
      if (tile >= TILES) break;
      const int mt = tile / NTL, nt = tile - mt * NTL;
      for (int kc = 0; kc < kch; ++kc) {
        const bf16x8 a = *(const bf16x8*)&sDY[mt * 16 + r][kc * 32 + q * 8];
        const bf16x8 bw = *(const bf16x8*)&sX[nt * 16 + r][kc * 32 + q * 8];
        acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw, acc[i], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < TPW; ++i) {
    const int tile = wave + 4 * i;
    if (tile >= TILES) break;
    const int mt = tile / NTL, nt = tile - mt * NTL;
    const int c = nt * 16 + r;
    if (c >= CIN) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = mt * 16 + q * 4 + j;
      part[(((long long)split * 27 + tap) * COUT + n) * CIN + c] = acc[i][j];
    }
  }
}

// dW[n][tap*CIN + c] = sum over splits (fixed order) of part[split][tap][n][c].
__global__ void k_fold_wgrad(const float* __restrict__ part, int nsplit, int cin, int cout,
                             float* __restrict__ dW) {
  const long long total = 27LL * cout * cin;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int c = (int)(e % cin);
  const long long t2 = e / cin;
  const int n = (int)(t2 % cout);
  const int tap = (int)(t2 / cout);
  float s = 0.f;
  for (int sp = 0; sp < nsplit; ++sp) s += part[(((long long)sp * 27 + tap) * cout + n) * cin + c];
  dW[(long long)n * 27 * cin + tap * cin + c] = s;
}

template <int CIN, int COUT>
static void wgrad_t(const Vol& v, const __bf16* dY, const __bf16* X, int xs, int xo, float* part,
                    int nsplit, float* dW, hipStream_t s) {
  hipLaunchKernelGGL((k_conv3_wgrad<CIN, COUT>), dim3(nsplit, 27), dim3(256), 0, s, v, dY, X, xs, xo, part, nsplit);
  const long long total = 27LL * CIN * COUT;
  hipLaunchKernelGGL(k_fold_wgrad, dim3((int)((total + 255) / 256)), dim3(256), 0, s, part, nsplit, CIN, COUT, dW);
}

void conv3_wgrad(int cin, int cout, const Vol& v, const __bf16* dY, const __bf16* X, int xs, int xo,
                 float* part, int nsplit, float* dW, hipStream_t s) {
#define AQU_WG(CI, CO) \
  if (cin == CI && cout == CO) return wgrad_t<CI, CO>(v, dY, X, xs, xo, part, nsplit, dW, s);
  AQU_WG(8, 16) AQU_WG(16, 16) AQU_WG(16, 32) AQU_WG(32, 32) AQU_WG(32, 64) AQU_WG(64, 64)
  AQU_WG(96, 32) AQU_WG(48, 16)
#undef AQU_WG
}

// ------------------------------------------------------------------ ReLU backward + bias gradient
// dYpre = dY * (Y > 0) as bf16 (dense, C channels), and per-block per-channel sums
// of dYpre (bias gradient partials, folded by k_fold_rows). Thread t works on
// channel t % C, so each channel's partial is summed in a fixed order.
__global__ void k_relu_bwd(long long V, int C, const __bf16* __restrict__ Y, int ysd, int yo,
                           const float* __restrict__ dY, int dys, int dyo, bool relu,
                           __bf16* __restrict__ dYpre, float* __restrict__ bpart, int vox_per_block) {
  __shared__ float red[256];
  const int c = threadIdx.x % C;
  const int lanesPerC = 256 / C;
  const int slot = threadIdx.x / C;
  float acc = 0.f;
  const long long v0 = (long long)blockIdx.x * vox_per_block;
  if (slot < lanesPerC) {
    for (int i = slot; i < vox_per_block; i += lanesPerC) {
      const long long vv = v0 + i;
      if (vv >= V) break;
      float g = dY[vv * dys + dyo + c];
      if (relu && !(bf2f(Y[vv * ysd + yo + c]) > 0.f)) g = 0.f;
      const __bf16 gb = f2bf(g);
      dYpre[vv * C + c] = gb;
      acc += bf2f(gb);
    }
  }
  red[threadIdx.x] = (slot < lanesPerC) ? acc : 0.f;
  __syncthreads();
  if (threadIdx.x < C) {
    float s = 0.f;
    for (int k = 0; k < lanesPerC; ++k) s += red[k * C + threadIdx.x];
    bpart[(long long)blockIdx.x * C + threadIdx.x] = s;
  }
}

// out[c] = sum over rows (fixed order) of part[row][c].
__global__ void k_fold_rows(const float* __restrict__ part, int rows, int C, float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  float s = 0.f;
  for (int r = 0; r < rows; ++r) s += part[(long long)r * C + c];
  out[c] = s;
}

void relu_bwd(long long V, int C, const __bf16* Y, int ys, int yo, const float* dY, int dys, int dyo,
              bool relu, __bf16* dYpre, float* bpart, int nblk, float* dbias, hipStream_t s) {
  const int vpb = (int)((V + nblk - 1) / nblk);
  hipLaunchKernelGGL(k_relu_bwd, dim3(nblk), dim3(256), 0, s, V, C, Y, ys, yo, dY, dys, dyo, relu,
                     dYpre, bpart, vpb);
  hipLaunchKernelGGL(k_fold_rows, dim3((C + 63) / 64), dim3(64), 0, s, bpart, nblk, C, dbias);
}

// ------------------------------------------------------------------ pooling / upsampling
// 2x2x2 average pooling (forward, bf16) and its adjoint (accumulating FP32).
// One thread per output voxel and 8 consecutive channels; the eight children are
// summed in the same (d, h, w) order for every channel.
__global__ void k_avgpool(Vol vo, int C, const __bf16* __restrict__ X, int xs, int xo,
                          __bf16* __restrict__ Y, int row0) {
  const int C8 = C >> 3;
  const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (unsigned)vo.W * C8) return;
  const int w = (int)(e / C8), j = (int)(e - (unsigned)w * C8);
  unsigned row = row0 + blockIdx.y;  // (b, d, h) of the output
  const int h = (int)(row % vo.H);
  row /= vo.H;
  const int d = (int)(row % vo.D), b = (int)(row / vo.D);
  const Vol vi{vo.B, vo.D * 2, vo.H * 2, vo.W * 2};
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const bf16x8 x8 =
        *(const bf16x8*)(X + vox_id(vi, b, 2 * d + (k >> 2), 2 * h + ((k >> 1) & 1), 2 * w + (k & 1)) * xs + xo + 8 * j);
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] += bf2f(x8[c]);
  }
  bf16x8 o8;
#pragma unroll
  for (int c = 0; c < 8; ++c) o8[c] = f2bf(acc[c] * 0.125f);
  *(bf16x8*)(Y + vox_id(vo, b, d, h, w) * C + 8 * j) = o8;
}
__global__ void k_avgpool_bwd(Vol vo, int C, const float* __restrict__ dY, float* __restrict__ dX,
                              int dxs, int dxo) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long V = (long long)vo.B * vo.D * vo.H * vo.W;
  if (e >= V * C) return;
  const int c = (int)(e % C);
  const VoxelPos p = vox_of(vo, e / C);
  const Vol vi{vo.B, vo.D * 2, vo.H * 2, vo.W * 2};
  const float g = dY[e] * 0.125f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    float* o = dX + vox_id(vi, p.b, 2 * p.d + (k >> 2), 2 * p.h + ((k >> 1) & 1), 2 * p.w + (k & 1)) * dxs + dxo + c;
    *o += g;
  }
}
// Nearest 2x upsampling into a channel slice (forward) and its adjoint (sum of the
// eight children, FP32).
// One thread per output voxel and 8 consecutive channels (a 16-byte load and store).
__global__ void k_upsample(Vol vi, int C, const __bf16* __restrict__ X, __bf16* __restrict__ Y, int ys,
                           int yo, int row0) {
  const int C8 = C >> 3;
  const int Wo = vi.W * 2, Ho = vi.H * 2, Do = vi.D * 2;
  const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned nrow = (unsigned)Wo * C8;
  if (e >= nrow) return;
  const int w = (int)(e / C8), j = (int)(e - (unsigned)w * C8);
  unsigned row = row0 + blockIdx.y;  // (b, d, h) of the output
  const int h = (int)(row % Ho);
  row /= Ho;
  const int d = (int)(row % Do), b = (int)(row / Do);
  const long long o = (((long long)b * Do + d) * Ho + h) * Wo + w;
  const long long i = vox_id(vi, b, d >> 1, h >> 1, w >> 1);
  *(bf16x8*)(Y + o * ys + yo + 8 * j) = *(const bf16x8*)(X + i * C + 8 * j);
}
__global__ void k_upsample_bwd(Vol vi, int C, const float* __restrict__ dY, int dys, int dyo,
                               float* __restrict__ dX) {
  const Vol vo{vi.B, vi.D * 2, vi.H * 2, vi.W * 2};
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long V = (long long)vi.B * vi.D * vi.H * vi.W;
  if (e >= V * C) return;
  const int c = (int)(e % C);
  const VoxelPos p = vox_of(vi, e / C);
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    s += dY[vox_id(vo, p.b, 2 * p.d + (k >> 2), 2 * p.h + ((k >> 1) & 1), 2 * p.w + (k & 1)) * dys + dyo + c];
  dX[e] = s;
}

static inline int nb1(long long n) { return (int)((n + 255) / 256); }
void avgpool(const Vol& vo, int C, const __bf16* X, int xs, int xo, __bf16* Y, hipStream_t s) {
  const int rows = vo.B * vo.D * vo.H;
  for (int r0 = 0; r0 < rows; r0 += 65535) {  // grid.y limit
    const dim3 grid((unsigned)((vo.W * (C / 8) + 255) / 256), (unsigned)std::min(65535, rows - r0));
    hipLaunchKernelGGL(k_avgpool, grid, dim3(256), 0, s, vo, C, X, xs, xo, Y, r0);
  }
}
void avgpool_bwd(const Vol& vo, int C, const float* dY, float* dX, int dxs, int dxo, hipStream_t s) {
  const long long n = (long long)vo.B * vo.D * vo.H * vo.W * C;
  hipLaunchKernelGGL(k_avgpool_bwd, dim3(nb1(n)), dim3(256), 0, s, vo, C, dY, dX, dxs, dxo);
}
void upsample(const Vol& vi, int C, const __bf16* X, __bf16* Y, int ys, int yo, hipStream_t s) {
  const int rows = vi.B * vi.D * 2 * vi.H * 2;
  for (int r0 = 0; r0 < rows; r0 += 65535) {  // grid.y limit
    const dim3 grid((unsigned)((vi.W * 2 * (C / 8) + 255) / 256), (unsigned)std::min(65535, rows - r0));
    hipLaunchKernelGGL(k_upsample, grid, dim3(256), 0, s, vi, C, X, Y, ys, yo, r0);
  }
}
void upsample_bwd(const Vol& vi, int C, const float* dY, int dys, int dyo, float* dX, hipStream_t s) {
  const long long n = (long long)vi.B * vi.D * vi.H * vi.W * C;
  hipLaunchKernelGGL(k_upsample_bwd, dim3(nb1(n)), dim3(256), 0, s, vi, C, dY, dys, dyo, dX);
}

// ------------------------------------------------------------------ output head (1x1x1, C -> 1)
__global__ void k_head(long long V, int C, const __bf16* __restrict__ X, const float* __restrict__ w,
                       const float* __restrict__ b, float* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  float s = b[0];
  for (int c0 = 0; c0 < C; c0 += 8) {  // C is a multiple of 8: 16-byte loads
    const bf16x8 x8 = *(const bf16x8*)(X + i * C + c0);
#pragma unroll
    for (int c = 0; c < 8; ++c) s += w[c0 + c] * bf2f(x8[c]);
  }
  y[i] = s;
}
__global__ void k_head_bwd(long long V, int C, const __bf16* __restrict__ X, const float* __restrict__ w,
                           const float* __restrict__ dy, float* __restrict__ dX, float* __restrict__ part,
                           int vpb) {
  __shared__ float red[256 * 17];
  float acc[17];
  for (int c = 0; c <= C; ++c) acc[c] = 0.f;
  const long long v0 = (long long)blockIdx.x * vpb;
  for (int i = threadIdx.x; i < vpb; i += 256) {
    const long long vv = v0 + i;
    if (vv >= V) break;
    const float g = dy[vv];
    for (int c = 0; c < C; ++c) {
      dX[vv * C + c] = g * w[c];
      acc[c] += g * bf2f(X[vv * C + c]);
    }
    acc[C] += g;
  }
  for (int c = 0; c <= C; ++c) red[c * 256 + threadIdx.x] = acc[c];
  __syncthreads();
  if ((int)threadIdx.x <= C) {
    float s = 0.f;
    for (int k = 0; k < 256; ++k) s += red[threadIdx.x * 256 + k];
    part[(long long)blockIdx.x * (C + 1) + threadIdx.x] = s;
  }
}
void head_fwd(long long V, int C, const __bf16* X, const float* w, const float* b, float* y, hipStream_t s) {
  hipLaunchKernelGGL(k_head, dim3(nb1(V)), dim3(256), 0, s, V, C, X, w, b, y);
}
void head_bwd(long long V, int C, const __bf16* X, const float* w, const float* dy, float* dX,
              float* part, int nblk, float* dwb, hipStream_t s) {
  const int vpb = (int)((V + nblk - 1) / nblk);
  hipLaunchKernelGGL(k_head_bwd, dim3(nblk), dim3(256), 0, s, V, C, X, w, dy, dX, part, vpb);
  hipLaunchKernelGGL(k_fold_rows, dim3(1), dim3(64), 0, s, part, nblk, C + 1, dwb);
}

// ------------------------------------------------------------------ loss
// Masked mean-squared error over fluid voxels: per-block partials of sum (y-t)^2*m
// and sum m, folded in order; the gradient dy = 2 (y - t) m / sum m needs the
// total, so it is a second pass.
__global__ void k_loss_partial(long long V, const float* __restrict__ y, const float* __restrict__ t,
                               const float* __restrict__ m, double* __restrict__ part, int vpb) {
  __shared__ double s1[256], s2[256];
  double a = 0, c = 0;
  const long long v0 = (long long)blockIdx.x * vpb;
  for (int i = threadIdx.x; i < vpb; i += 256) {
    const long long vv = v0 + i;
    if (vv >= V) break;
    const double d = (double)y[vv] - (double)t[vv];
    a += d * d * m[vv];
    c += m[vv];
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = c;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
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
__global__ void k_loss_fold(const double* __restrict__ part, int n, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double a = 0, c = 0;
  for (int i = 0; i < n; ++i) {
    a += part[2 * i];
    c += part[2 * i + 1];
  }
  out[0] = c > 0 ? a / c : 0.0;
  out[1] = c;
}
__global__ void k_loss_grad(long long V, const float* __restrict__ y, const float* __restrict__ t,
                            const float* __restrict__ m, const double* __restrict__ tot,
                            float* __restrict__ dy) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  const double c = tot[1];
  dy[i] = c > 0 ? (float)(2.0 * ((double)y[i] - (double)t[i]) * m[i] / c) : 0.f;
}
void masked_mse(long long V, const float* y, const float* t, const float* m, double* part, int nblk,
                double* out, float* dy, hipStream_t s) {
  const int vpb = (int)((V + nblk - 1) / nblk);
  hipLaunchKernelGGL(k_loss_partial, dim3(nblk), dim3(256), 0, s, V, y, t, m, part, vpb);
  hipLaunchKernelGGL(k_loss_fold, dim3(1), dim3(64), 0, s, part, nblk, out);
  if (dy) hipLaunchKernelGGL(k_loss_grad, dim3(nb1(V)), dim3(256), 0, s, V, y, t, m, out, dy);
}

// ------------------------------------------------------------------ parameters
// Adam on FP32 master weights (bias-corrected step sizes passed in), then the bf16
// copies the convolutions read.
__global__ void k_adam(long long n, float* __restrict__ w, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ vv, float lr_t, float b1, float b2,
                       float eps) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * vv[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  vv[i] = vi;
  w[i] -= lr_t * mi / (sqrtf(vi) + eps);
}
void adam(long long n, float* w, const float* g, float* m, float* v, float lr_t, float b1, float b2,
          float eps, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3(nb1(n)), dim3(256), 0, s, n, w, g, m, v, lr_t, b1, b2, eps);
}

// Forward weights Wk[n][tap*cin + c] and flipped transposed input-gradient weights
// Wt[c][tap'*cout + n] = W[n][(26 - tap')*cin + c], both bf16.
__global__ void k_pack_weights(int cin, int cout, const float* __restrict__ W, __bf16* __restrict__ Wk,
                               __bf16* __restrict__ Wt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = 27 * cin * cout;
  if (e >= total) return;
  const int c = e % cin;
  const int t2 = e / cin;
  const int tap = t2 % 27;
  const int n = t2 / 27;
  const __bf16 v = f2bf(W[e]);
  Wk[e] = v;
  if (Wt) Wt[(long long)c * 27 * cout + (26 - tap) * cout + n] = v;
}
void pack_weights(int cin, int cout, const float* W, __bf16* Wk, __bf16* Wt, hipStream_t s) {
  const int total = 27 * cin * cout;
  hipLaunchKernelGGL(k_pack_weights, dim3((total + 255) / 256), dim3(256), 0, s, cin, cout, W, Wk, Wt);
}

// ------------------------------------------------------------------ fills
__global__ void k_zero_f32(float* __restrict__ p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long st = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = 0.f;
}
void zero_f32(float* p, long long n, hipStream_t s) {
  int nb = (int)std::min<long long>((n + 255) / 256, 8192);
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_zero_f32, dim3(nb), dim3(256), 0, s, p, n);
}

// Probe of the MFMA lane layout assumed above: D = A (16x32) * B (32x16), A and B
// given row-major in FP32, D row-major.
__global__ void k_mfma_probe(const float* __restrict__ A, const float* __restrict__ B, float* __restrict__ D) {
  const int l = threadIdx.x;
  const int r = l & 15, q = l >> 4;
  bf16x8 a, b;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    a[j] = f2bf(A[r * 32 + q * 8 + j]);
    b[j] = f2bf(B[(q * 8 + j) * 16 + r]);
  }
  f32x4 c = {0.f, 0.f, 0.f, 0.f};
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
#pragma unroll
  for (int i = 0; i < 4; ++i) D[(q * 4 + i) * 16 + r] = c[i];
}
void mfma_probe(const float* A, const float* B, float* D, hipStream_t s) {
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, A, B, D);
}

}  // namespace aqu

// ------------------------------------------------------------------ FLIP grid <-> network tensors
// The network sees the FLIP interior as one volume with (D, H, W) = (X, Y, Z), the
// same x-major, z-fastest order as the cell hash.
#include "aqflip_internal.h"

namespace aqu {

__device__ __forceinline__ void cell3(const aqf::Geom& G, long long n, int& x, int& y, int& z) {
  const uint32_t u = (uint32_t)n;  // n < X*Y*Z < 2^31
  x = (int)aqf::fast_div(u, G.divYZ);
  const uint32_t r = u - (uint32_t)x * (uint32_t)(G.Y * G.Z);
  y = (int)aqf::fast_div(r, G.divZ);
  z = (int)(r - (uint32_t)y * (uint32_t)G.Z);
}

// sigma = RMS of the right-hand side b over FLUID cells: sig[0] = sigma, sig[1] = 1/sigma
// (0 when there is no fluid or b vanishes).
__global__ void k_nn_sigma_part(aqf::Geom G, const uint8_t* __restrict__ content,
                                const float* __restrict__ rhs, double* __restrict__ part) {
  __shared__ double s1[256], s2[256];
  double a = 0, c = 0;
  const long long n0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long st = (long long)gridDim.x * blockDim.x;
  for (long long n = n0; n < G.ncells_hash; n += st) {
    int x, y, z;
    cell3(G, n, x, y, z);
    const long long gi = G.g(x, y, z);
    if (content[gi] != aqf::FLUID) continue;
    const double b = (double)rhs[gi];
    a += b * b;
    c += 1.0;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = c;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
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
// Sums of the two columns of n partial pairs, by one 256-thread block: thread t adds
// pairs t, t + 256, ... in order, then a fixed-shape tree (deterministic). The result
// is returned to thread 0.
__device__ void fold_pairs(const double* __restrict__ part, int n, double& a, double& c) {
  __shared__ double s1[256], s2[256];
  double x = 0, y = 0;
  for (int i = (int)threadIdx.x; i < n; i += 256) {
    x += part[2 * i];
    y += part[2 * i + 1];
  }
  s1[threadIdx.x] = x;
  s2[threadIdx.x] = y;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  a = s1[0];
  c = s2[0];
}
__global__ void k_nn_sigma_fold(const double* __restrict__ part, int n, double* __restrict__ sig) {
  double a, c;
  fold_pairs(part, n, a, c);
  if (threadIdx.x != 0) return;
  const double s = (c > 0) ? sqrt(a / c) : 0.0;
  sig[0] = s;
  sig[1] = s > 0 ? 1.0 / s : 0.0;
}
void nn_sigma(const aqf::Geom& G, const uint8_t* content, const float* rhs, double* part, int nparts,
              double* sig, hipStream_t s) {
  hipLaunchKernelGGL(k_nn_sigma_part, dim3(nparts), dim3(256), 0, s, G, content, rhs, part);
  hipLaunchKernelGGL(k_nn_sigma_fold, dim3(1), dim3(256), 0, s, part, nparts, sig);
}

__global__ void k_nn_input(aqf::Geom G, const uint8_t* __restrict__ content, const float* __restrict__ rhs,
                           const double* __restrict__ sig, __bf16* __restrict__ x0, float* __restrict__ mask) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.ncells_hash) return;
  int x, y, z;
  cell3(G, n, x, y, z);
  const long long gi = G.g(x, y, z);
  const bool fl = content[gi] == aqf::FLUID;
  const float b = fl ? (float)((double)rhs[gi] * sig[1]) : 0.f;
  __bf16* o = x0 + n * 8;
  o[0] = f2bf(b);
  o[1] = f2bf(fl ? 1.f : 0.f);
#pragma unroll
  for (int c = 2; c < 8; ++c) o[c] = f2bf(0.f);
  if (mask) mask[n] = fl ? 1.f : 0.f;
}
void nn_input(const aqf::Geom& G, const uint8_t* content, const float* rhs, const double* sig,
              __bf16* x0, float* mask, hipStream_t s) {
  hipLaunchKernelGGL(k_nn_input, dim3((G.ncells_hash + 255) / 256), dim3(256), 0, s, G, content, rhs, sig,
                     x0, mask);
}

__global__ void k_nn_target(aqf::Geom G, const uint8_t* __restrict__ content, const float* __restrict__ p,
                            const double* __restrict__ sig, float scale, float* __restrict__ t) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.ncells_hash) return;
  int x, y, z;
  cell3(G, n, x, y, z);
  const long long gi = G.g(x, y, z);
  const bool fl = content[gi] == aqf::FLUID;
  t[n] = (fl && sig[0] > 0) ? (float)((double)p[gi] / (sig[0] * (double)scale)) : 0.f;
}
void nn_target(const aqf::Geom& G, const uint8_t* content, const float* p, const double* sig, float scale,
               float* t, hipStream_t s) {
  hipLaunchKernelGGL(k_nn_target, dim3((G.ncells_hash + 255) / 256), dim3(256), 0, s, G, content, p, sig,
                     scale, t);
}

__global__ void k_nn_output(aqf::Geom G, const uint8_t* __restrict__ content, const float* __restrict__ y,
                            const double* __restrict__ sig, float scale, float* __restrict__ p) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.ncells_hash) return;
  int x, yy, z;
  cell3(G, n, x, yy, z);
  const long long gi = G.g(x, yy, z);
  const bool fl = content[gi] == aqf::FLUID;
  p[gi] = fl ? (float)((double)y[n] * sig[0] * (double)scale) : 0.f;
}
void nn_output(const aqf::Geom& G, const uint8_t* content, const float* y, const double* sig, float scale,
               float* p, hipStream_t s) {
  hipLaunchKernelGGL(k_nn_output, dim3((G.ncells_hash + 255) / 256), dim3(256), 0, s, G, content, y, sig,
                     scale, p);
}

// ------------------------------------------------------------------ warm-start guess
// The guess runs on dense fields in hash order (x-major, z-fastest). On grids
// larger than the training resolution the right-hand side is restricted 2x per
// level (sum of the eight children; a coarse cell is FLUID when at least four of
// its children are), the network runs on the coarsest level, and its pressure is
// prolonged back with cell-centred trilinear weights (3/4, 1/4 per axis). The
// fine guess p is finally scaled by alpha = <p, b> / <p, A p> (A = the pressure
// operator of the FP64 residual, symmetric positive definite), the alpha that
// minimises the energy-norm error ||p* - alpha p||_A with A p* = b: whatever the
// network's scale error, the scaled guess is never further from the solution in
// that norm than a zero guess. (Minimising the residual norm instead would weigh
// the guess's high-frequency error, which the first sweeps remove anyway, and
// shrink its useful smooth part.)
__global__ void k_nn_gather(aqf::Geom G, const uint8_t* __restrict__ content, const float* __restrict__ rhs,
                            float* __restrict__ b, float* __restrict__ m) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.ncells_hash) return;
  int x, y, z;
  cell3(G, n, x, y, z);
  const long long gi = G.g(x, y, z);
  const bool fl = content[gi] == aqf::FLUID;
  b[n] = fl ? rhs[gi] : 0.f;
  m[n] = fl ? 1.f : 0.f;
}
__global__ void k_nn_restrict(int Xc, int Yc, int Zc, const float* __restrict__ bf, const float* __restrict__ mf,
                              float* __restrict__ bc, float* __restrict__ mc) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (long long)Xc * Yc * Zc) return;
  const int z = (int)(n % Zc);
  const long long t = n / Zc;
  const int y = (int)(t % Yc), x = (int)(t / Yc);
  const int Yf = 2 * Yc, Zf = 2 * Zc;
  float sb = 0.f, sm = 0.f;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const long long f = ((long long)(2 * x + (i >> 2)) * Yf + (2 * y + ((i >> 1) & 1))) * Zf + 2 * z + (i & 1);
    sb += bf[f];
    sm += mf[f];
  }
  const bool fl = sm >= 4.f;
  bc[n] = fl ? sb : 0.f;
  mc[n] = fl ? 1.f : 0.f;
}
__global__ void k_nn_sigma_dense(long long V, const float* __restrict__ b, const float* __restrict__ m,
                                 double* __restrict__ part) {
  __shared__ double s1[256], s2[256];
  double a = 0, c = 0;
  for (long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x; n < V;
       n += (long long)gridDim.x * blockDim.x) {
    if (m[n] == 0.f) continue;
    const double v = (double)b[n];
    a += v * v;
    c += 1.0;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = c;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
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
__global__ void k_nn_input_dense(long long V, const float* __restrict__ b, const float* __restrict__ m,
                                 const double* __restrict__ sig, __bf16* __restrict__ x0) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= V) return;
  __bf16* o = x0 + n * 8;
  o[0] = f2bf((float)((double)b[n] * sig[1]));
  o[1] = f2bf(m[n]);
#pragma unroll
  for (int c = 2; c < 8; ++c) o[c] = f2bf(0.f);
}
__global__ void k_nn_output_dense(long long V, const float* __restrict__ y, const float* __restrict__ m,
                                  const double* __restrict__ sig, float scale, float* __restrict__ p) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= V) return;
  p[n] = m[n] != 0.f ? (float)((double)y[n] * sig[0] * (double)scale) : 0.f;
}
__global__ void k_nn_prolong(int Xc, int Yc, int Zc, const float* __restrict__ pc, float* __restrict__ pf) {
  const int Xf = 2 * Xc, Yf = 2 * Yc, Zf = 2 * Zc;
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (long long)Xf * Yf * Zf) return;
  const int z = (int)(n % Zf);
  const long long t = n / Zf;
  const int y = (int)(t % Yf), x = (int)(t / Yf);
  // parent and the neighbouring parent on the child's side, weights 3/4 and 1/4
  const int px = x >> 1, py = y >> 1, pz = z >> 1;
  const int qx = min(max(px + ((x & 1) ? 1 : -1), 0), Xc - 1);
  const int qy = min(max(py + ((y & 1) ? 1 : -1), 0), Yc - 1);
  const int qz = min(max(pz + ((z & 1) ? 1 : -1), 0), Zc - 1);
  const int ix[2] = {px, qx}, iy[2] = {py, qy}, iz[2] = {pz, qz};
  const float w[2] = {0.75f, 0.25f};
  float s = 0.f;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c)
        s += w[a] * w[b] * w[c] * pc[((long long)ix[a] * Yc + iy[b]) * Zc + iz[c]];
  pf[n] = s;
}
__global__ void k_nn_scatter(aqf::Geom G, const uint8_t* __restrict__ content, const float* __restrict__ pf,
                             float* __restrict__ p) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.ncells_hash) return;
  int x, y, z;
  cell3(G, n, x, y, z);
  const long long gi = G.g(x, y, z);
  p[gi] = content[gi] == aqf::FLUID ? pf[n] : 0.f;
}
// <p, b> and <p, A p> over FLUID cells, A as in the FP64 residual.
// Per block: sum of p*b and p*(A p) over FLUID cells in a grid-stride loop, four
// cells per trip whose loads (content and p at the cell and its six neighbours, b)
// are all issued before any is used; each thread adds its cells in stride order.
__global__ void k_nn_alpha_part(aqf::Geom G, const uint8_t* __restrict__ content, const float* __restrict__ rhs,
                                const float* __restrict__ p, double* __restrict__ part) {
  __shared__ double s1[256], s2[256];
  double qb = 0, qq = 0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long o[6] = {G.sx, -G.sx, (long long)G.EZ, -(long long)G.EZ, 1, -1};
  for (long long n0 = (long long)blockIdx.x * blockDim.x + threadIdx.x; n0 < G.ncells_hash; n0 += 4 * stride) {
    uint8_t cc[4], cn[4][6];
    float pc[4], pn[4][6], bc[4];
    bool ok[4], in[4][6];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const long long n = n0 + k * stride;
      ok[k] = n < G.ncells_hash;
      int x, y, z;
      cell3(G, ok[k] ? n : 0, x, y, z);
      const long long gi = G.g(x, y, z);
      in[k][0] = x + 1 < G.X, in[k][1] = x > 0, in[k][2] = y + 1 < G.Y;
      in[k][3] = y > 0, in[k][4] = z + 1 < G.Z, in[k][5] = z > 0;
      cc[k] = content[gi];
      pc[k] = p[gi];
      bc[k] = rhs[gi];
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const long long gn = in[k][j] ? gi + o[j] : gi;
        cn[k][j] = content[gn];
        pn[k][j] = p[gn];
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (!ok[k] || cc[k] != aqf::FLUID) continue;
      const double q = (double)pc[k];
      double ap = 6.0 * q;
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        if (!in[k][j])
          ap -= q;
        else if (cn[k][j] == aqf::FLUID)
          ap -= (double)pn[k][j];
      }
      qb += q * (double)bc[k];
      qq += q * ap;
    }
  }
  s1[threadIdx.x] = qb;
  s2[threadIdx.x] = qq;
  __syncthreads();
  for (int r = 128; r > 0; r >>= 1) {
    if ((int)threadIdx.x < r) {
      s1[threadIdx.x] += s1[threadIdx.x + r];
      s2[threadIdx.x] += s2[threadIdx.x + r];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = s2[0];
  }
}
__global__ void k_nn_alpha_fold(const double* __restrict__ part, int n, double* __restrict__ alpha) {
  double qb, qq;
  fold_pairs(part, n, qb, qq);
  if (threadIdx.x != 0) return;
  alpha[0] = (qq > 0 && isfinite(qb / qq)) ? qb / qq : 0.0;
}
__global__ void k_nn_alpha_apply(aqf::Geom G, const double* __restrict__ alpha, float* __restrict__ p) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.ncells_hash) return;
  int x, y, z;
  cell3(G, n, x, y, z);
  const long long gi = G.g(x, y, z);
  p[gi] = (float)((double)p[gi] * alpha[0]);
}

// r = b - A p over FLUID cells (FP64 arithmetic, stored FP32), 0 elsewhere.
__global__ void k_nn_residual(aqf::Geom G, const uint8_t* __restrict__ content, const float* __restrict__ rhs,
                              const float* __restrict__ p, float* __restrict__ r) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.ncells_hash) return;
  int x, y, z;
  cell3(G, n, x, y, z);
  const long long gi = G.g(x, y, z);
  if (content[gi] != aqf::FLUID) {
    r[gi] = 0.f;
    return;
  }
  const double pc = (double)p[gi];
  double ap = 6.0 * pc;
  const bool in[6] = {x + 1 < G.X, x > 0, y + 1 < G.Y, y > 0, z + 1 < G.Z, z > 0};
  const long long o[6] = {G.sx, -G.sx, (long long)G.EZ, -(long long)G.EZ, 1, -1};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    if (!in[k])
      ap -= pc;
    else if (content[gi + o[k]] == aqf::FLUID)
      ap -= (double)p[gi + o[k]];
  }
  r[gi] = (float)((double)rhs[gi] - ap);
}
// p += d on FLUID cells.
__global__ void k_nn_axpy(aqf::Geom G, const uint8_t* __restrict__ content, const float* __restrict__ d,
                          float* __restrict__ p) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.ncells_hash) return;
  int x, y, z;
  cell3(G, n, x, y, z);
  const long long gi = G.g(x, y, z);
  if (content[gi] == aqf::FLUID) p[gi] += d[gi];
}

static inline int nblk(long long n) { return (int)((n + 255) / 256); }

void nn_residual(const aqf::Geom& G, const uint8_t* content, const float* rhs, const float* p, float* r,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_nn_residual, dim3(nblk(G.ncells_hash)), dim3(256), 0, s, G, content, rhs, p, r);
}
void nn_axpy(const aqf::Geom& G, const uint8_t* content, const float* d, float* p, hipStream_t s) {
  hipLaunchKernelGGL(k_nn_axpy, dim3(nblk(G.ncells_hash)), dim3(256), 0, s, G, content, d, p);
}

void nn_guess_gather(const aqf::Geom& G, const uint8_t* content, const float* rhs, float* b, float* m,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_nn_gather, dim3(nblk(G.ncells_hash)), dim3(256), 0, s, G, content, rhs, b, m);
}
void nn_restrict(int Xc, int Yc, int Zc, const float* bf, const float* mf, float* bc, float* mc, hipStream_t s) {
  hipLaunchKernelGGL(k_nn_restrict, dim3(nblk((long long)Xc * Yc * Zc)), dim3(256), 0, s, Xc, Yc, Zc, bf, mf,
                     bc, mc);
}
void nn_input_dense(long long V, const float* b, const float* m, double* part, int nparts, double* sig,
                    __bf16* x0, hipStream_t s) {
  hipLaunchKernelGGL(k_nn_sigma_dense, dim3(nparts), dim3(256), 0, s, V, b, m, part);
  hipLaunchKernelGGL(k_nn_sigma_fold, dim3(1), dim3(256), 0, s, part, nparts, sig);
  hipLaunchKernelGGL(k_nn_input_dense, dim3(nblk(V)), dim3(256), 0, s, V, b, m, sig, x0);
}
void nn_output_dense(long long V, const float* y, const float* m, const double* sig, float scale, float* p,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_nn_output_dense, dim3(nblk(V)), dim3(256), 0, s, V, y, m, sig, scale, p);
}
void nn_prolong(int Xc, int Yc, int Zc, const float* pc, float* pf, hipStream_t s) {
  hipLaunchKernelGGL(k_nn_prolong, dim3(nblk(8LL * Xc * Yc * Zc)), dim3(256), 0, s, Xc, Yc, Zc, pc, pf);
}
void nn_scatter_scaled(const aqf::Geom& G, const uint8_t* content, const float* rhs, const float* pf, float* p,
                       double* part, int nparts, double* alpha, hipStream_t s) {
  hipLaunchKernelGGL(k_nn_scatter, dim3(nblk(G.ncells_hash)), dim3(256), 0, s, G, content, pf, p);
  hipLaunchKernelGGL(k_nn_alpha_part, dim3(nparts), dim3(256), 0, s, G, content, rhs, p, part);
  hipLaunchKernelGGL(k_nn_alpha_fold, dim3(1), dim3(256), 0, s, part, nparts, alpha);
  hipLaunchKernelGGL(k_nn_alpha_apply, dim3(nblk(G.ncells_hash)), dim3(256), 0, s, G, alpha, p);
}

// Host tensor (V x 2 floats) -> network input (V x 8 bf16).
__global__ void k_host_input(long long V, const float* __restrict__ x2, __bf16* __restrict__ x0) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= V) return;
  __bf16* o = x0 + n * 8;
  o[0] = f2bf(x2[2 * n]);
  o[1] = f2bf(x2[2 * n + 1]);
#pragma unroll
  for (int c = 2; c < 8; ++c) o[c] = f2bf(0.f);
}
void host_input(long long V, const float* x2, __bf16* x0, hipStream_t s) {
  hipLaunchKernelGGL(k_host_input, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, V, x2, x0);
}

}  // namespace aqu
