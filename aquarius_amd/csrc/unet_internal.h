// unet_internal.h -- kernels of the warm-start U-Net (unet_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aqu {

struct Vol {
  int B, D, H, W;  // batch and spatial extent; voxel id ((b*D + d)*H + h)*W + w
};

void conv3_forward(int cin, int cout, const Vol& v, const __bf16* X, int xs, int xo, const __bf16* Wk,
                   const float* bias, __bf16* Y, int ys, int yo, bool relu, hipStream_t s);
void conv3_dgrad(int cin, int cout, const Vol& v, const __bf16* dY, const __bf16* Wt, float* dX,
                 int dxs, int dxo, hipStream_t s);
void conv3_wgrad(int cin, int cout, const Vol& v, const __bf16* dY, const __bf16* X, int xs, int xo,
                 float* part, int nsplit, float* dW, hipStream_t s);
void relu_bwd(long long V, int C, const __bf16* Y, int ys, int yo, const float* dY, int dys, int dyo,
              bool relu, __bf16* dYpre, float* bpart, int nblk, float* dbias, hipStream_t s);
void avgpool(const Vol& vo, int C, const __bf16* X, int xs, int xo, __bf16* Y, hipStream_t s);
void avgpool_bwd(const Vol& vo, int C, const float* dY, float* dX, int dxs, int dxo, hipStream_t s);
void upsample(const Vol& vi, int C, const __bf16* X, __bf16* Y, int ys, int yo, hipStream_t s);
void upsample_bwd(const Vol& vi, int C, const float* dY, int dys, int dyo, float* dX, hipStream_t s);
void head_fwd(long long V, int C, const __bf16* X, const float* w, const float* b, float* y, hipStream_t s);
void head_bwd(long long V, int C, const __bf16* X, const float* w, const float* dy, float* dX,
              float* part, int nblk, float* dwb, hipStream_t s);
void masked_mse(long long V, const float* y, const float* t, const float* m, double* part, int nblk,
                double* out, float* dy, hipStream_t s);
void adam(long long n, float* w, const float* g, float* m, float* v, float lr_t, float b1, float b2,
          float eps, hipStream_t s);
void pack_weights(int cin, int cout, const float* W, __bf16* Wk, __bf16* Wt, hipStream_t s);
void zero_f32(float* p, long long n, hipStream_t s);
void mfma_probe(const float* A, const float* B, float* D, hipStream_t s);

}  // namespace aqu
