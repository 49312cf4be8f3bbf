// unet_api.cpp -- host side of the warm-start U-Net (include/aqflip_unet.h).
//
// Network (channels-last bf16 activations, FP32 accumulation, FP32 master weights):
//   level 1 (D):   c1a 8->16, c1b 16->16 (written into cat1[0:16])
//   level 2 (D/2): avgpool, c2a 16->32, c2b 32->32 (into cat2[0:32])
//   level 3 (D/4): avgpool, c3a 32->64, c3b 64->64
//   up to D/2:     nearest upsample into cat2[32:96], c4a 96->32, c4b 32->32
//   up to D:       nearest upsample into cat1[16:48], c5a 48->16, c5b 16->16
//   head:          1x1x1 16->1 (FP32, no activation)
// Every 3x3x3 convolution is followed by ReLU. Input channel 0 is b/sigma, channel
// 1 the FLUID mask, channels 2..7 zero (padding to the 8-channel MFMA k-slice).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/aqflip_unet.h"
#include "aqflip_internal.h"
#include "flip_view.h"
#include "unet_internal.h"

using namespace aqu;

namespace {

aq_status ufail(aq_status s, const std::string& m) {
  aqf::set_last_error(m);
  return s;
}

#define UHIP(expr)                                                                              \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      return ufail(e_ == hipErrorOutOfMemory ? AQ_ENOMEM : AQ_EHIP,                            \
                   std::string(#expr) + ": " + hipGetErrorString(e_));                          \
  } while (0)
#define UTRY(expr)               \
  do {                           \
    aq_status s_ = (expr);       \
    if (s_ != AQ_OK) return s_;  \
  } while (0)

struct ConvSpec {
  int cin, cout;
};
// c1a c1b c2a c2b c3a c3b c4a c4b c5a c5b
const ConvSpec kConvs[10] = {{8, 16},  {16, 16}, {16, 32}, {32, 32}, {32, 64},
                             {64, 64}, {96, 32}, {32, 32}, {48, 16}, {16, 16}};
constexpr int kHeadC = 16;
constexpr int kNSplit = 256;  // weight-gradient voxel splits
constexpr int kRedBlocks = 512;

}  // namespace

struct aq_unet {
  aq_unet_config cfg{};
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;

  long long nparam = 0;
  long long woff[10]{}, boff[10]{}, head_w = 0, head_b = 0;
  float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
  __bf16* Wk[10]{};
  __bf16* Wt[10]{};
  long long adam_t = 0;

  // activations / gradients for the current capacity
  long long capV = 0;  // voxels at level 1 (B*D*H*W)
  __bf16 *x0 = nullptr, *a1 = nullptr, *cat1 = nullptr, *p1 = nullptr, *a2 = nullptr,
         *cat2 = nullptr, *p2 = nullptr, *a3 = nullptr, *b3 = nullptr, *a4 = nullptr,
         *b4 = nullptr, *a5 = nullptr, *b5 = nullptr, *dypre = nullptr;
  float *y = nullptr, *gy = nullptr, *g16a = nullptr, *g16b = nullptr, *gcat1 = nullptr,
        *g32a = nullptr, *g32b = nullptr, *gcat2 = nullptr, *g64a = nullptr, *g64b = nullptr,
        *gp1 = nullptr, *gp2 = nullptr, *target = nullptr, *mask = nullptr;
  float *wpart = nullptr, *bpart = nullptr;
  double *lpart = nullptr, *lout = nullptr, *sig = nullptr;
  std::vector<void*> act_allocs;

  // warm-start guess: dense per-level fields (level 0 = the FLIP grid), sized for
  // the largest grid seen
  long long guess_cap = 0;
  std::vector<float*> gb, gm, gp;
  float *res = nullptr, *corr = nullptr;  // residual and correction, FLIP grid layout
  long long res_cap = 0;
  std::vector<void*> guess_allocs, res_allocs;
  double* alpha = nullptr;

  // dataset
  int ds_cap = 0, ds_n = 0, ds_D = 0, ds_H = 0, ds_W = 0;
  __bf16* ds_x = nullptr;
  float *ds_t = nullptr, *ds_m = nullptr;

  ~aq_unet() {
    if (stream) hipStreamSynchronize(stream);
    for (void* p : guess_allocs) hipFree(p);
    for (void* p : res_allocs) hipFree(p);
    for (void* p : act_allocs) hipFree(p);
    for (void* p : allocs) hipFree(p);
    if (stream) hipStreamDestroy(stream);
  }
  template <typename T>
  aq_status alloc(T** p, size_t n, std::vector<void*>& list) {
    UHIP(hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)));
    list.push_back(*p);
    UHIP(hipMemsetAsync(*p, 0, std::max<size_t>(n, 1) * sizeof(T), stream));
    return AQ_OK;
  }

  // Dense guess fields for a grid of V cells and `levels` restrictions.
  aq_status ensure_guess(long long V, int levels) {
    if (V <= guess_cap && (int)gb.size() > levels) return AQ_OK;
    UHIP(hipStreamSynchronize(stream));
    for (void* p : guess_allocs) hipFree(p);
    guess_allocs.clear();
    gb.assign(levels + 1, nullptr);
    gm.assign(levels + 1, nullptr);
    gp.assign(levels + 1, nullptr);
    for (int l = 0; l <= levels; ++l) {
      const long long Vl = (V >> (3 * l)) + 1;
      UTRY(alloc(&gb[l], Vl, guess_allocs));
      UTRY(alloc(&gm[l], Vl, guess_allocs));
      UTRY(alloc(&gp[l], Vl, guess_allocs));
    }
    guess_cap = V;
    return AQ_OK;
  }

  // Residual and correction fields in the FLIP grid's padded layout (ncell floats).
  aq_status ensure_res(long long ncell) {
    if (ncell <= res_cap) return AQ_OK;
    UHIP(hipStreamSynchronize(stream));
    for (void* p : res_allocs) hipFree(p);
    res_allocs.clear();
    UTRY(alloc(&res, ncell, res_allocs));
    UTRY(alloc(&corr, ncell, res_allocs));
    res_cap = ncell;
    return AQ_OK;
  }

  aq_status ensure(long long V) {
    if (V <= capV) return AQ_OK;
    UHIP(hipStreamSynchronize(stream));
    for (void* p : act_allocs) hipFree(p);
    act_allocs.clear();
    const long long V2 = V / 8 + 1, V3 = V / 64 + 1;
    auto& L = act_allocs;
    UTRY(alloc(&x0, V * 8, L));
    UTRY(alloc(&a1, V * 16, L));
    UTRY(alloc(&cat1, V * 48, L));
    UTRY(alloc(&p1, V2 * 16, L));
    UTRY(alloc(&a2, V2 * 32, L));
    UTRY(alloc(&cat2, V2 * 96, L));
    UTRY(alloc(&p2, V3 * 32, L));
    UTRY(alloc(&a3, V3 * 64, L));
    UTRY(alloc(&b3, V3 * 64, L));
    UTRY(alloc(&a4, V2 * 32, L));
    UTRY(alloc(&b4, V2 * 32, L));
    UTRY(alloc(&a5, V * 16, L));
    UTRY(alloc(&b5, V * 16, L));
    UTRY(alloc(&dypre, V * 64, L));
    UTRY(alloc(&y, V, L));
    UTRY(alloc(&gy, V, L));
    UTRY(alloc(&g16a, V * 16, L));
    UTRY(alloc(&g16b, V * 16, L));
    UTRY(alloc(&gcat1, V * 48, L));
    UTRY(alloc(&g32a, V2 * 32, L));
    UTRY(alloc(&g32b, V2 * 32, L));
    UTRY(alloc(&gcat2, V2 * 96, L));
    UTRY(alloc(&g64a, V3 * 64, L));
    UTRY(alloc(&g64b, V3 * 64, L));
    UTRY(alloc(&gp1, V2 * 16, L));
    UTRY(alloc(&gp2, V3 * 32, L));
    UTRY(alloc(&target, V, L));
    UTRY(alloc(&mask, V, L));
    capV = V;
    return AQ_OK;
  }

  void pack() {
    for (int l = 0; l < 10; ++l)
      pack_weights(kConvs[l].cin, kConvs[l].cout, params + woff[l], Wk[l], l == 0 ? nullptr : Wt[l],
                   stream);
  }

  // Forward pass over x0 (bf16, 8 channels) into y (FP32).
  void forward(const Vol& v1) {
    const Vol v2{v1.B, v1.D / 2, v1.H / 2, v1.W / 2}, v3{v1.B, v1.D / 4, v1.H / 4, v1.W / 4};
    auto conv = [&](int l, const Vol& v, const __bf16* X, int xs, int xo, __bf16* Y, int ys, int yo) {
      conv3_forward(kConvs[l].cin, kConvs[l].cout, v, X, xs, xo, Wk[l], params + boff[l], Y, ys, yo,
                    true, stream);
    };
    conv(0, v1, x0, 8, 0, a1, 16, 0);
    conv(1, v1, a1, 16, 0, cat1, 48, 0);
    avgpool(v2, 16, cat1, 48, 0, p1, stream);
    conv(2, v2, p1, 16, 0, a2, 32, 0);
    conv(3, v2, a2, 32, 0, cat2, 96, 0);
    avgpool(v3, 32, cat2, 96, 0, p2, stream);
    conv(4, v3, p2, 32, 0, a3, 64, 0);
    conv(5, v3, a3, 64, 0, b3, 64, 0);
    upsample(v3, 64, b3, cat2, 96, 32, stream);
    conv(6, v2, cat2, 96, 0, a4, 32, 0);
    conv(7, v2, a4, 32, 0, b4, 32, 0);
    upsample(v2, 32, b4, cat1, 48, 16, stream);
    conv(8, v1, cat1, 48, 0, a5, 16, 0);
    conv(9, v1, a5, 16, 0, b5, 16, 0);
    const long long V = (long long)v1.B * v1.D * v1.H * v1.W;
    head_fwd(V, kHeadC, b5, params + head_w, params + head_b, y, stream);
  }

  // Backward pass from gy (dL/dy) into grads (every parameter overwritten).
  void backward(const Vol& v1) {
    const Vol v2{v1.B, v1.D / 2, v1.H / 2, v1.W / 2}, v3{v1.B, v1.D / 4, v1.H / 4, v1.W / 4};
    const long long V1 = (long long)v1.B * v1.D * v1.H * v1.W, V2 = V1 / 8, V3 = V1 / 64;
    // layer l: output Y (stride ys, offset yo) with gradient G (stride gs, offset go), input X.
    auto layer = [&](int l, const Vol& v, long long Vn, const __bf16* Y, int ys, int yo,
                     const float* G, int gs, int go, const __bf16* X, int xs, int xo, float* dX,
                     int dxs, int dxo) {
      const int ci = kConvs[l].cin, co = kConvs[l].cout;
      relu_bwd(Vn, co, Y, ys, yo, G, gs, go, true, dypre, bpart, kRedBlocks, grads + boff[l], stream);
      conv3_wgrad(ci, co, v, dypre, X, xs, xo, wpart, kNSplit, grads + woff[l], stream);
      if (dX) conv3_dgrad(ci, co, v, dypre, Wt[l], dX, dxs, dxo, stream);
    };
    head_bwd(V1, kHeadC, b5, params + head_w, gy, g16a, bpart, kRedBlocks, grads + head_w, stream);
    layer(9, v1, V1, b5, 16, 0, g16a, 16, 0, a5, 16, 0, g16b, 16, 0);        // c5b
    layer(8, v1, V1, a5, 16, 0, g16b, 16, 0, cat1, 48, 0, gcat1, 48, 0);    // c5a
    upsample_bwd(v2, 32, gcat1, 48, 16, g32a, stream);                       // -> d b4
    layer(7, v2, V2, b4, 32, 0, g32a, 32, 0, a4, 32, 0, g32b, 32, 0);        // c4b
    layer(6, v2, V2, a4, 32, 0, g32b, 32, 0, cat2, 96, 0, gcat2, 96, 0);    // c4a
    upsample_bwd(v3, 64, gcat2, 96, 32, g64a, stream);                       // -> d b3
    layer(5, v3, V3, b3, 64, 0, g64a, 64, 0, a3, 64, 0, g64b, 64, 0);        // c3b
    layer(4, v3, V3, a3, 64, 0, g64b, 64, 0, p2, 32, 0, gp2, 32, 0);        // c3a
    avgpool_bwd(v3, 32, gp2, gcat2, 96, 0, stream);                          // += d s2
    layer(3, v2, V2, cat2, 96, 0, gcat2, 96, 0, a2, 32, 0, g32a, 32, 0);    // c2b
    layer(2, v2, V2, a2, 32, 0, g32a, 32, 0, p1, 16, 0, gp1, 16, 0);        // c2a
    avgpool_bwd(v2, 16, gp1, gcat1, 48, 0, stream);                          // += d s1
    layer(1, v1, V1, cat1, 48, 0, gcat1, 48, 0, a1, 16, 0, g16a, 16, 0);    // c1b
    layer(0, v1, V1, a1, 16, 0, g16a, 16, 0, x0, 8, 0, nullptr, 0, 0);      // c1a
  }

  void adam_step() {
    ++adam_t;
    const double b1 = cfg.beta1, b2 = cfg.beta2;
    const double lr_t = cfg.lr * std::sqrt(1.0 - std::pow(b2, (double)adam_t)) /
                        (1.0 - std::pow(b1, (double)adam_t));
    adam(nparam, params, grads, m, v, (float)lr_t, cfg.beta1, cfg.beta2, cfg.eps, stream);
    pack();
  }
};

namespace aqu {
// Kernels of the FLIP <-> U-Net glue (unet_kernels.hip).
void nn_sigma(const aqf::Geom& G, const uint8_t* content, const float* rhs, double* part, int nparts,
              double* sig, hipStream_t s);
void nn_input(const aqf::Geom& G, const uint8_t* content, const float* rhs, const double* sig,
              __bf16* x0, float* mask, hipStream_t s);
void nn_target(const aqf::Geom& G, const uint8_t* content, const float* p, const double* sig,
               float scale, float* t, hipStream_t s);
void nn_output(const aqf::Geom& G, const uint8_t* content, const float* y, const double* sig,
               float scale, float* p, hipStream_t s);
void host_input(long long V, const float* x2, __bf16* x0, hipStream_t s);
void nn_guess_gather(const aqf::Geom& G, const uint8_t* content, const float* rhs, float* b, float* m,
                     hipStream_t s);
void nn_restrict(int Xc, int Yc, int Zc, const float* bf, const float* mf, float* bc, float* mc,
                 hipStream_t s);
void nn_input_dense(long long V, const float* b, const float* m, double* part, int nparts, double* sig,
                    __bf16* x0, hipStream_t s);
void nn_output_dense(long long V, const float* y, const float* m, const double* sig, float scale, float* p,
                     hipStream_t s);
void nn_prolong(int Xc, int Yc, int Zc, const float* pc, float* pf, hipStream_t s);
void nn_scatter_scaled(const aqf::Geom& G, const uint8_t* content, const float* rhs, const float* pf,
                       float* p, double* part, int nparts, double* alpha, hipStream_t s);
void nn_residual(const aqf::Geom& G, const uint8_t* content, const float* rhs, const float* p, float* r,
                 hipStream_t s);
void nn_axpy(const aqf::Geom& G, const uint8_t* content, const float* d, float* p, hipStream_t s);
}  // namespace aqu

namespace {

float out_scale(const aq_unet* net, int D) {
  const double r = (double)D / 64.0;
  return (float)(net->cfg.target_scale * r * r);
}

// dst = alpha * net(src) on FLUID cells (0 elsewhere) for the FLIP grid of fv, where
// src is a right-hand side in the grid layout (the pressure right-hand side, or a
// residual) and alpha = <d, src> / <d, A d> for the network's output d. Grids larger
// than cfg.infer_extent are restricted 2x per level for the network and its output
// prolonged back. Runs on the network's stream; the caller synchronises.
aq_status nn_infer(const aqf::FlipView& fv, aq_unet* net, const float* src, float* dst) {
  const aqf::Geom& G = fv.G;
  int levels = 0;  // halve while larger than the network's resolution and still divisible
  while (std::max(G.X, std::max(G.Y, G.Z)) >> levels > net->cfg.infer_extent &&
         (G.X >> levels) % 8 == 0 && (G.Y >> levels) % 8 == 0 && (G.Z >> levels) % 8 == 0)
    ++levels;
  const int X = G.X >> levels, Y = G.Y >> levels, Z = G.Z >> levels;
  if (X % 4 || Y % 4 || Z % 4) return ufail(AQ_EINVAL, "grid extents must be multiples of 4");
  const long long V0 = (long long)G.X * G.Y * G.Z, V = (long long)X * Y * Z;
  UTRY(net->ensure(V));
  UTRY(net->ensure_guess(V0, levels));
  hipStream_t s = net->stream;
  nn_guess_gather(G, fv.gr->content, src, net->gb[0], net->gm[0], s);
  for (int l = 1; l <= levels; ++l)
    nn_restrict(G.X >> l, G.Y >> l, G.Z >> l, net->gb[l - 1], net->gm[l - 1], net->gb[l], net->gm[l], s);
  nn_input_dense(V, net->gb[levels], net->gm[levels], net->lpart, kRedBlocks, net->sig, net->x0, s);
  net->forward(Vol{1, X, Y, Z});
  nn_output_dense(V, net->y, net->gm[levels], net->sig, out_scale(net, X), net->gp[levels], s);
  for (int l = levels; l >= 1; --l)
    nn_prolong(G.X >> l, G.Y >> l, G.Z >> l, net->gp[l], net->gp[l - 1], s);
  nn_scatter_scaled(G, fv.gr->content, src, net->gp[0], dst, net->lpart, kRedBlocks, net->alpha, s);
  UHIP(hipGetLastError());
  return AQ_OK;
}

aq_status upload_input(aq_unet* net, const float* x, long long V) {
  // x (V*2 floats) -> x0 (bf16, 8 channels) through a staging copy in `y`/`gy`
  float* stage = net->g16b;  // V*16 floats of scratch, enough for V*2
  UHIP(hipMemcpyAsync(stage, x, V * 2 * sizeof(float), hipMemcpyHostToDevice, net->stream));
  host_input(V, stage, net->x0, net->stream);
  return AQ_OK;
}

}  // namespace

extern "C" {

aq_status aq_unet_config_default(aq_unet_config* c) {
  if (!c) return ufail(AQ_EINVAL, "null config");
  std::memset(c, 0, sizeof(*c));
  c->channels[0] = 16;
  c->channels[1] = 32;
  c->channels[2] = 64;
  c->seed = 1;
  c->lr = 1e-3f;
  c->beta1 = 0.9f;
  c->beta2 = 0.999f;
  c->eps = 1e-8f;
  c->target_scale = 64.f;
  c->infer_extent = 64;
  return AQ_OK;
}

aq_status aq_unet_create(const aq_unet_config* cfg, aq_unet** out) {
  if (!cfg || !out) return ufail(AQ_EINVAL, "null argument");
  if (cfg->channels[0] != 16 || cfg->channels[1] != 32 || cfg->channels[2] != 64)
    return ufail(AQ_EINVAL, "the U-Net kernels are built for 16/32/64 channels");
  if (!(cfg->lr > 0) || !(cfg->target_scale > 0)) return ufail(AQ_EINVAL, "lr and target_scale must be > 0");
  if (cfg->infer_extent < 4) return ufail(AQ_EINVAL, "infer_extent must be >= 4");
  aq_unet* n = new aq_unet;
  n->cfg = *cfg;
  auto bail = [&](aq_status s) {
    delete n;
    return s;
  };
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(ufail(AQ_EHIP, "stream creation failed"));
  long long off = 0;
  for (int l = 0; l < 10; ++l) {
    n->woff[l] = off;
    off += 27LL * kConvs[l].cin * kConvs[l].cout;
    n->boff[l] = off;
    off += kConvs[l].cout;
  }
  n->head_w = off;
  off += kHeadC;
  n->head_b = off;
  off += 1;
  n->nparam = off;
  aq_status st = AQ_OK;
  if (st == AQ_OK) st = n->alloc(&n->params, off, n->allocs);
  if (st == AQ_OK) st = n->alloc(&n->grads, off, n->allocs);
  if (st == AQ_OK) st = n->alloc(&n->m, off, n->allocs);
  if (st == AQ_OK) st = n->alloc(&n->v, off, n->allocs);
  for (int l = 0; l < 10 && st == AQ_OK; ++l) {
    const size_t nw = 27ull * kConvs[l].cin * kConvs[l].cout;
    st = n->alloc(&n->Wk[l], nw, n->allocs);
    if (st == AQ_OK) st = n->alloc(&n->Wt[l], nw, n->allocs);
  }
  size_t maxw = 0;
  for (int l = 0; l < 10; ++l) maxw = std::max<size_t>(maxw, 27ull * kConvs[l].cin * kConvs[l].cout);
  if (st == AQ_OK) st = n->alloc(&n->wpart, maxw * kNSplit, n->allocs);
  if (st == AQ_OK) st = n->alloc(&n->bpart, (size_t)kRedBlocks * 64, n->allocs);
  if (st == AQ_OK) st = n->alloc(&n->lpart, (size_t)kRedBlocks * 2, n->allocs);
  if (st == AQ_OK) st = n->alloc(&n->alpha, 1, n->allocs);
  if (st == AQ_OK) st = n->alloc(&n->lout, 4, n->allocs);
  if (st == AQ_OK) st = n->alloc(&n->sig, 4, n->allocs);
  if (st != AQ_OK) return bail(st);
  // He-normal initialisation over the real fan-in (the first layer sees 2 channels).
  std::vector<float> w((size_t)off, 0.f);
  std::mt19937_64 rng(cfg->seed);
  for (int l = 0; l < 10; ++l) {
    const int ci = kConvs[l].cin, co = kConvs[l].cout;
    const int real_ci = l == 0 ? 2 : ci;
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (27.f * real_ci)));
    for (int o = 0; o < co; ++o)
      for (int t = 0; t < 27; ++t)
        for (int c = 0; c < ci; ++c)
          w[n->woff[l] + ((size_t)o * 27 + t) * ci + c] = c < real_ci ? nd(rng) : 0.f;
  }
  std::normal_distribution<float> nh(0.f, std::sqrt(1.f / kHeadC));
  for (int c = 0; c < kHeadC; ++c) w[n->head_w + c] = nh(rng);
  aq_status s2 = aq_unet_set_params(n, w.data());
  if (s2 != AQ_OK) return bail(s2);
  *out = n;
  return AQ_OK;
}

void aq_unet_destroy(aq_unet* net) { delete net; }

int64_t aq_unet_param_count(const aq_unet* net) { return net ? net->nparam : 0; }

aq_status aq_unet_get_params(aq_unet* net, float* host) {
  if (!net || !host) return ufail(AQ_EINVAL, "null argument");
  UHIP(hipMemcpyAsync(host, net->params, net->nparam * sizeof(float), hipMemcpyDeviceToHost, net->stream));
  UHIP(hipStreamSynchronize(net->stream));
  return AQ_OK;
}

aq_status aq_unet_set_params(aq_unet* net, const float* host) {
  if (!net || !host) return ufail(AQ_EINVAL, "null argument");
  UHIP(hipMemcpyAsync(net->params, host, net->nparam * sizeof(float), hipMemcpyHostToDevice, net->stream));
  net->pack();
  UHIP(hipStreamSynchronize(net->stream));
  return AQ_OK;
}

aq_status aq_unet_adam_step_host(aq_unet* net, const float* grads) {
  if (!net || !grads) return ufail(AQ_EINVAL, "null argument");
  UHIP(hipMemcpyAsync(net->grads, grads, net->nparam * sizeof(float), hipMemcpyHostToDevice, net->stream));
  net->adam_step();
  UHIP(hipGetLastError());
  UHIP(hipStreamSynchronize(net->stream));
  return AQ_OK;
}

aq_status aq_unet_set_lr(aq_unet* net, float lr) {
  if (!net || !(lr >= 0.f) || !std::isfinite(lr)) return ufail(AQ_EINVAL, "lr must be finite and >= 0");
  net->cfg.lr = lr;
  return AQ_OK;
}

aq_status aq_unet_get_adam_state(aq_unet* net, float* m, float* v, int64_t* t) {
  if (!net) return ufail(AQ_EINVAL, "null argument");
  if (m) UHIP(hipMemcpyAsync(m, net->m, net->nparam * sizeof(float), hipMemcpyDeviceToHost, net->stream));
  if (v) UHIP(hipMemcpyAsync(v, net->v, net->nparam * sizeof(float), hipMemcpyDeviceToHost, net->stream));
  UHIP(hipStreamSynchronize(net->stream));
  if (t) *t = net->adam_t;
  return AQ_OK;
}

aq_status aq_unet_forward_host(aq_unet* net, const float* x, int32_t B, int32_t D, int32_t H, int32_t W,
                               float* yh) {
  if (!net || !x || !yh) return ufail(AQ_EINVAL, "null argument");
  if (B < 1 || D < 4 || H < 4 || W < 4 || D % 4 || H % 4 || W % 4)
    return ufail(AQ_EINVAL, "extents must be multiples of 4");
  const long long V = (long long)B * D * H * W;
  UTRY(net->ensure(V));
  UTRY(upload_input(net, x, V));
  net->forward(Vol{B, D, H, W});
  UHIP(hipGetLastError());
  UHIP(hipMemcpyAsync(yh, net->y, V * sizeof(float), hipMemcpyDeviceToHost, net->stream));
  UHIP(hipStreamSynchronize(net->stream));
  return AQ_OK;
}

aq_status aq_unet_gradients_host(aq_unet* net, const float* x, const float* t, const float* mk,
                                 int32_t B, int32_t D, int32_t H, int32_t W, double* loss,
                                 float* gh) {
  if (!net || !x || !t || !mk) return ufail(AQ_EINVAL, "null argument");
  if (B < 1 || D < 4 || H < 4 || W < 4 || D % 4 || H % 4 || W % 4 || W > 64)
    return ufail(AQ_EINVAL, "extents must be multiples of 4 with W <= 64");
  const long long V = (long long)B * D * H * W;
  UTRY(net->ensure(V));
  UTRY(upload_input(net, x, V));
  UHIP(hipMemcpyAsync(net->target, t, V * sizeof(float), hipMemcpyHostToDevice, net->stream));
  UHIP(hipMemcpyAsync(net->mask, mk, V * sizeof(float), hipMemcpyHostToDevice, net->stream));
  const Vol v1{B, D, H, W};
  net->forward(v1);
  masked_mse(V, net->y, net->target, net->mask, net->lpart, kRedBlocks, net->lout, net->gy, net->stream);
  net->backward(v1);
  UHIP(hipGetLastError());
  double lo[2];
  UHIP(hipMemcpyAsync(lo, net->lout, sizeof(lo), hipMemcpyDeviceToHost, net->stream));
  if (gh) UHIP(hipMemcpyAsync(gh, net->grads, net->nparam * sizeof(float), hipMemcpyDeviceToHost, net->stream));
  UHIP(hipStreamSynchronize(net->stream));
  if (loss) *loss = lo[0];
  return AQ_OK;
}

aq_status aq_unet_dataset_reserve(aq_unet* net, int32_t cap, int32_t D, int32_t H, int32_t W) {
  if (!net || cap < 1) return ufail(AQ_EINVAL, "bad argument");
  if (D % 4 || H % 4 || W % 4 || W > 64) return ufail(AQ_EINVAL, "extents must be multiples of 4, W <= 64");
  UHIP(hipStreamSynchronize(net->stream));
  if (net->ds_x) {
    for (void* p : {(void*)net->ds_x, (void*)net->ds_t, (void*)net->ds_m}) {
      hipFree(p);
      net->allocs.erase(std::remove(net->allocs.begin(), net->allocs.end(), p), net->allocs.end());
    }
  }
  const long long V = (long long)D * H * W;
  UTRY(net->alloc(&net->ds_x, (size_t)cap * V * 8, net->allocs));
  UTRY(net->alloc(&net->ds_t, (size_t)cap * V, net->allocs));
  UTRY(net->alloc(&net->ds_m, (size_t)cap * V, net->allocs));
  net->ds_cap = cap;
  net->ds_n = 0;
  net->ds_D = D, net->ds_H = H, net->ds_W = W;
  UHIP(hipStreamSynchronize(net->stream));
  return AQ_OK;
}

aq_status aq_unet_dataset_add_flip(aq_unet* net, aq_flip* flip) {
  if (!net || !flip) return ufail(AQ_EINVAL, "null argument");
  if (net->ds_n >= net->ds_cap) return ufail(AQ_EINVAL, "dataset is full");
  aqf::FlipView fv = aqf::flip_view(flip);
  if (fv.G.X != net->ds_D || fv.G.Y != net->ds_H || fv.G.Z != net->ds_W)
    return ufail(AQ_EINVAL, "grid extent differs from the dataset's");
  UHIP(hipStreamSynchronize(fv.stream));
  const long long V = (long long)net->ds_D * net->ds_H * net->ds_W;
  const size_t k = (size_t)net->ds_n;
  nn_sigma(fv.G, fv.gr->content, fv.gr->rhs, net->lpart, kRedBlocks, net->sig, net->stream);
  nn_input(fv.G, fv.gr->content, fv.gr->rhs, net->sig, net->ds_x + k * V * 8, net->ds_m + k * V, net->stream);
  nn_target(fv.G, fv.gr->content, fv.gr->p[*fv.pb], net->sig, out_scale(net, fv.G.X), net->ds_t + k * V,
            net->stream);
  UHIP(hipGetLastError());
  UHIP(hipStreamSynchronize(net->stream));
  net->ds_n++;
  return AQ_OK;
}

int32_t aq_unet_dataset_size(const aq_unet* net) { return net ? net->ds_n : 0; }

aq_status aq_unet_train(aq_unet* net, int32_t steps, int32_t batch, int32_t train_count, uint64_t seed,
                        float* hist) {
  if (!net || steps < 0 || batch < 1) return ufail(AQ_EINVAL, "bad argument");
  if (train_count < 1 || train_count > net->ds_n) return ufail(AQ_EINVAL, "train_count out of range");
  const long long V = (long long)net->ds_D * net->ds_H * net->ds_W;
  UTRY(net->ensure(V * batch));
  const Vol v1{batch, net->ds_D, net->ds_H, net->ds_W};
  std::mt19937_64 rng(seed);
  std::uniform_int_distribution<int> pick(0, train_count - 1);
  std::vector<double> lo((size_t)std::max(steps, 1) * 2);
  for (int s = 0; s < steps; ++s) {
    for (int b = 0; b < batch; ++b) {
      const size_t k = (size_t)pick(rng);
      UHIP(hipMemcpyAsync(net->x0 + b * V * 8, net->ds_x + k * V * 8, V * 8 * sizeof(__bf16),
                          hipMemcpyDeviceToDevice, net->stream));
      UHIP(hipMemcpyAsync(net->target + b * V, net->ds_t + k * V, V * sizeof(float),
                          hipMemcpyDeviceToDevice, net->stream));
      UHIP(hipMemcpyAsync(net->mask + b * V, net->ds_m + k * V, V * sizeof(float),
                          hipMemcpyDeviceToDevice, net->stream));
    }
    net->forward(v1);
    masked_mse(V * batch, net->y, net->target, net->mask, net->lpart, kRedBlocks, net->lout, net->gy,
               net->stream);
    UHIP(hipMemcpyAsync(&lo[2 * s], net->lout, 2 * sizeof(double), hipMemcpyDeviceToHost, net->stream));
    net->backward(v1);
    net->adam_step();
    UHIP(hipGetLastError());
  }
  UHIP(hipStreamSynchronize(net->stream));
  if (hist)
    for (int s = 0; s < steps; ++s) hist[s] = (float)lo[2 * s];
  return AQ_OK;
}

aq_status aq_unet_eval(aq_unet* net, int32_t first, int32_t count, double* loss) {
  if (!net || !loss || first < 0 || count < 1 || first + count > net->ds_n)
    return ufail(AQ_EINVAL, "bad argument");
  const long long V = (long long)net->ds_D * net->ds_H * net->ds_W;
  UTRY(net->ensure(V));
  const Vol v1{1, net->ds_D, net->ds_H, net->ds_W};
  double num = 0, den = 0;
  for (int i = first; i < first + count; ++i) {
    UHIP(hipMemcpyAsync(net->x0, net->ds_x + (size_t)i * V * 8, V * 8 * sizeof(__bf16),
                        hipMemcpyDeviceToDevice, net->stream));
    net->forward(v1);
    masked_mse(V, net->y, net->ds_t + (size_t)i * V, net->ds_m + (size_t)i * V, net->lpart, kRedBlocks,
               net->lout, nullptr, net->stream);
    double lo[2];
    UHIP(hipMemcpyAsync(lo, net->lout, sizeof(lo), hipMemcpyDeviceToHost, net->stream));
    UHIP(hipStreamSynchronize(net->stream));
    num += lo[0] * lo[1];
    den += lo[1];
  }
  *loss = den > 0 ? num / den : 0.0;
  return AQ_OK;
}

aq_status aq_flip_nn_guess(aq_flip* flip, aq_unet* net) {
  if (!net || !flip) return ufail(AQ_EINVAL, "null argument");
  aqf::FlipView fv = aqf::flip_view(flip);
  // The flip stream's work (divergence) must finish before the net reads it, and the
  // guess must land before the flip stream continues: both streams meet here.
  UHIP(hipStreamSynchronize(fv.stream));
  UTRY(nn_infer(fv, net, fv.gr->rhs, fv.gr->p[*fv.pb]));
  UHIP(hipStreamSynchronize(net->stream));
  return AQ_OK;
}

aq_status aq_flip_nn_correct(aq_flip* flip, aq_unet* net) {
  if (!net || !flip) return ufail(AQ_EINVAL, "null argument");
  aqf::FlipView fv = aqf::flip_view(flip);
  UTRY(net->ensure_res(fv.G.ncell));
  UHIP(hipStreamSynchronize(fv.stream));
  float* p = fv.gr->p[*fv.pb];
  nn_residual(fv.G, fv.gr->content, fv.gr->rhs, p, net->res, net->stream);
  UTRY(nn_infer(fv, net, net->res, net->corr));
  nn_axpy(fv.G, fv.gr->content, net->corr, p, net->stream);
  UHIP(hipGetLastError());
  UHIP(hipStreamSynchronize(net->stream));
  return AQ_OK;
}

aq_status aq_unet_mfma_probe(const float* A, const float* B, float* D) {
  if (!A || !B || !D) return ufail(AQ_EINVAL, "null argument");
  float *dA, *dB, *dD;
  UHIP(hipMalloc((void**)&dA, 16 * 32 * sizeof(float)));
  UHIP(hipMalloc((void**)&dB, 32 * 16 * sizeof(float)));
  UHIP(hipMalloc((void**)&dD, 16 * 16 * sizeof(float)));
  UHIP(hipMemcpy(dA, A, 16 * 32 * sizeof(float), hipMemcpyHostToDevice));
  UHIP(hipMemcpy(dB, B, 32 * 16 * sizeof(float), hipMemcpyHostToDevice));
  mfma_probe(dA, dB, dD, nullptr);
  UHIP(hipGetLastError());
  UHIP(hipMemcpy(D, dD, 16 * 16 * sizeof(float), hipMemcpyDeviceToHost));
  hipFree(dA);
  hipFree(dB);
  hipFree(dD);
  return AQ_OK;
}

}  // extern "C"
