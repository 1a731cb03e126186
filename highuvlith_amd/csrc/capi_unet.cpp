// C ABI of the U-Net warm start (include/huv.h): a 2-D U-Net with `levels`
// resolution levels of base * 2^l channels (two 3x3 conv + ReLU per level and
// decoder stage, 2x2 max pooling, nearest 2x upsampling with skip
// concatenation, 1x1 FP32 head), BF16 weights and activations with FP32
// accumulation on the matrix cores, FP32 master parameters and Adam.  It maps
// a target pattern to an initial ILT theta; there is no reference
// counterpart (the reference starts ILT from the target, ilt.rs:495-522).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/huv.h"
#include "capi_internal.hpp"
#include "hip/unet_kernels.hpp"

using namespace huvi;
using huvk::bf16_t;

namespace {

struct Conv {
  int cin = 0, cout = 0;
  int64_t w_off = 0, b_off = 0;  // offsets into the flat FP32 parameters
};

// Device tensors of one (N, H, W) shape.
struct Acts {
  int N = 0, H = 0, W = 0;
  std::vector<void*> owned;
  bf16_t* x8 = nullptr;
  std::vector<bf16_t*> e1, e2, pool, cat, d1, d2;        // forward
  std::vector<bf16_t*> ge1, ge2, gpool, gcat, gd1, gd2;  // gradients
  bf16_t *xt = nullptr, *dyt = nullptr, *tmp = nullptr;
  float *out = nullptr, *dout = nullptr, *part = nullptr, *hpart = nullptr, *bpart = nullptr;
  double* lpart = nullptr;
  size_t part_floats = 0;
  void release() {
    for (void* p : owned) hipFree(p);
    owned.clear();
  }
};

constexpr int kLossBlocks = 256;
constexpr int kHeadBlocks = 256;

}  // namespace

struct huv_unet {
  huv_ctx* ctx = nullptr;
  int levels = 0, base = 0;
  std::vector<Conv> enc1, enc2, dec1, dec2;
  int64_t head_w = 0, head_b = 0, nparams = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  bf16_t *wb = nullptr, *wr = nullptr;  // BF16 copy of all parameters; rotated conv weights
  int64_t t = 0;
  Acts a;
  ~huv_unet() {
    a.release();
    for (void* q : {(void*)p, (void*)g, (void*)m, (void*)v, (void*)wb, (void*)wr})
      if (q) hipFree(q);
  }
  int ch(int l) const { return base << l; }
};

namespace {

hipStream_t S(huv_unet* u) { return u->ctx->stream(); }

#define HCHK(expr, what)                          \
  do {                                            \
    hipError_t e_ = (expr);                       \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

huv_status alloc(Acts& a, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes ? bytes : 256);
  if (e != hipSuccess) return hip_fail(e, "U-Net activation allocation");
  a.owned.push_back(*p);
  return HUV_OK;
}

template <class T>
huv_status alloc_t(Acts& a, T** p, size_t count) {
  return alloc(a, (void**)p, count * sizeof(T));
}

// The weight gradient reads the NHWC tensors directly where the row width
// allows 64-pixel segments; otherwise it reads NCHW copies.
bool wgrad_nhwc(const Conv& c, int W) { return W % 64 == 0 && c.cin % 8 == 0 && c.cout % 8 == 0; }

int wgrad_splits(const Conv& c, int N, int H, int W) {
  const int rows = N * H;
  if (wgrad_nhwc(c, W)) {
    // about 1024 work-groups (four 3-wave groups per CU)
    const int tiles = ((c.cout + 31) / 32) * ((c.cin + 31) / 32);
    return std::max(1, std::min((1024 + tiles - 1) / tiles, rows));
  }
  const int cob = (c.cout + 31) / 32;
  const int tiles = (cob >= 2 ? (cob + 1) / 2 : cob) * ((c.cin + 31) / 32) * 3;
  int s = (1024 + tiles - 1) / tiles;
  s = std::min(s, std::max(1, rows / 4));
  return std::max(1, s);
}

// Chunks of the bias-gradient reduction: about 2048 work-groups in total,
// at least 4096 pixels each.
int bias_splits(const Conv& c, int N, int H, int W) {
  const int64_t total = (int64_t)N * H * W;
  int64_t s = std::max<int64_t>(1, 2048 / c.cout);
  s = std::min<int64_t>(s, std::max<int64_t>(1, total / 4096));
  return (int)s;
}

huv_status ensure_acts(huv_unet* u, int N, int H, int W) {
  Acts& a = u->a;
  if (a.N == N && a.H == H && a.W == W) return HUV_OK;
  hipDeviceSynchronize();
  a.release();
  a = Acts{};
  a.N = N;
  a.H = H;
  a.W = W;
  const int L = u->levels;
  const int64_t P = (int64_t)N * H * W;
  huv_status st;
#define A(ptr, count) if ((st = alloc_t(a, &(ptr), (size_t)(count))) != HUV_OK) return st
  A(a.x8, P * 8);
  a.e1.resize(L); a.e2.resize(L); a.ge1.resize(L); a.ge2.resize(L);
  a.pool.resize(L); a.gpool.resize(L); a.cat.resize(L); a.gcat.resize(L);
  a.d1.resize(L); a.d2.resize(L); a.gd1.resize(L); a.gd2.resize(L);
  size_t maxt = (size_t)P * 8;
  size_t part = 0, bpart = 0;
  auto note_conv = [&](const Conv& c, int l) {
    const int Hl = H >> l, Wl = W >> l;
    const size_t ws = (size_t)wgrad_splits(c, N, Hl, Wl);
    part = std::max(part, ws * c.cout * 9 * c.cin);
    part = std::max(part, (size_t)bias_splits(c, N, Hl, Wl) * c.cout);
    bpart = std::max(bpart, ws * c.cout);
  };
  for (int l = 0; l < L; ++l) {
    const int64_t Pl = P >> (2 * l);
    const int c = u->ch(l);
    A(a.e1[l], Pl * c); A(a.e2[l], Pl * c); A(a.ge1[l], Pl * c); A(a.ge2[l], Pl * c);
    maxt = std::max(maxt, (size_t)(Pl * c));
    note_conv(u->enc1[l], l);
    note_conv(u->enc2[l], l);
    if (l < L - 1) {
      A(a.pool[l], (Pl / 4) * c); A(a.gpool[l], (Pl / 4) * c);
      const int ct = u->ch(l + 1) + c;
      A(a.cat[l], Pl * ct); A(a.gcat[l], Pl * ct);
      A(a.d1[l], Pl * c); A(a.d2[l], Pl * c); A(a.gd1[l], Pl * c); A(a.gd2[l], Pl * c);
      maxt = std::max(maxt, (size_t)(Pl * ct));
      note_conv(u->dec1[l], l);
      note_conv(u->dec2[l], l);
    }
  }
  A(a.xt, maxt); A(a.dyt, maxt); A(a.tmp, maxt);
  A(a.out, P); A(a.dout, P);
  a.part_floats = std::max(part, (size_t)1);
  A(a.part, a.part_floats);
  A(a.bpart, std::max(bpart, (size_t)1));
  A(a.hpart, (size_t)kHeadBlocks * (u->base + 1));
  A(a.lpart, kLossBlocks);
#undef A
  return HUV_OK;
}

const bf16_t* wbf(huv_unet* u, const Conv& c) { return u->wb + c.w_off; }
const bf16_t* wrot(huv_unet* u, const Conv& c) { return u->wr + c.w_off; }

// BF16 weights (all parameters) and the rotated copies used by the data gradient.
huv_status refresh_bf16(huv_unet* u) {
  HCHK(huvk::launch_to_bf16(u->p, u->nparams, u->wb, S(u)), "U-Net bf16 cast");
  for (auto* list : {&u->enc1, &u->enc2, &u->dec1, &u->dec2})
    for (const Conv& c : *list)
      HCHK(huvk::launch_conv_weight_rot(wbf(u, c), c.cout, c.cin, u->wr + c.w_off, S(u)),
           "U-Net weight rotation");
  return HUV_OK;
}

huv_status conv_fwd(huv_unet* u, const Conv& c, const bf16_t* x, int l, bf16_t* y) {
  const Acts& a = u->a;
  HCHK(huvk::launch_conv3x3_fwd(x, a.N, a.H >> l, a.W >> l, c.cin, wbf(u, c), u->p + c.b_off,
                                c.cout, 1, y, S(u)),
       "U-Net conv forward");
  return HUV_OK;
}

huv_status forward(huv_unet* u, const float* in) {
  Acts& a = u->a;
  const int L = u->levels;
  const int64_t P = (int64_t)a.N * a.H * a.W;
  HCHK(huvk::launch_pack_input(in, P, a.x8, S(u)), "U-Net input");
  huv_status st;
  for (int l = 0; l < L; ++l) {
    const bf16_t* src = l == 0 ? a.x8 : a.pool[l - 1];
    if ((st = conv_fwd(u, u->enc1[l], src, l, a.e1[l])) != HUV_OK) return st;
    if ((st = conv_fwd(u, u->enc2[l], a.e1[l], l, a.e2[l])) != HUV_OK) return st;
    if (l < L - 1)
      HCHK(huvk::launch_maxpool_fwd(a.e2[l], a.N, a.H >> l, a.W >> l, u->ch(l), a.pool[l], S(u)),
           "U-Net pool");
  }
  for (int l = L - 2; l >= 0; --l) {
    const bf16_t* up = l == L - 2 ? a.e2[L - 1] : a.d2[l + 1];
    HCHK(huvk::launch_upcat_fwd(up, u->ch(l + 1), a.e2[l], u->ch(l), a.N, a.H >> l, a.W >> l,
                                a.cat[l], S(u)),
         "U-Net upsample");
    if ((st = conv_fwd(u, u->dec1[l], a.cat[l], l, a.d1[l])) != HUV_OK) return st;
    if ((st = conv_fwd(u, u->dec2[l], a.d1[l], l, a.d2[l])) != HUV_OK) return st;
  }
  HCHK(huvk::launch_head_fwd(a.d2[0], P, u->base, u->p + u->head_w, u->p + u->head_b, a.out,
                             S(u)),
       "U-Net head");
  return HUV_OK;
}

// Gradients of one conv (+ReLU) layer: masks dy in place, writes dW, db into
// u->g and (if dx) the input gradient.
huv_status conv_bwd(huv_unet* u, const Conv& c, const bf16_t* x, const bf16_t* y, bf16_t* dy,
                    int l, bf16_t* dx) {
  Acts& a = u->a;
  const int H = a.H >> l, W = a.W >> l;
  const int splits = wgrad_splits(c, a.N, H, W);
  if (wgrad_nhwc(c, W)) {
    if (y) HCHK(huvk::launch_relu_mask(dy, y, a.N, H, W, c.cout, S(u)), "U-Net relu mask");
    HCHK(huvk::launch_conv3x3_wgrad_nhwc(dy, x, a.N, H, W, c.cin, c.cout, splits, a.part, a.bpart,
                                         S(u)),
         "U-Net weight gradient");
    HCHK(huvk::launch_sum_splits(a.part, splits, (int64_t)c.cout * 9 * c.cin, u->g + c.w_off,
                                 S(u)),
         "U-Net split sum");
    HCHK(huvk::launch_sum_splits(a.bpart, splits, c.cout, u->g + c.b_off, S(u)),
         "U-Net bias split sum");
    if (dx)
      HCHK(huvk::launch_conv3x3_fwd(dy, a.N, H, W, c.cout, wrot(u, c), nullptr, c.cin, 0, dx,
                                    S(u)),
           "U-Net data gradient");
    return HUV_OK;
  }
  HCHK(huvk::launch_relu_mask_t(dy, y, a.N, H, W, c.cout, a.dyt, S(u)), "U-Net relu mask");
  HCHK(huvk::launch_nhwc_to_nchw(x, a.N, H, W, c.cin, a.xt, S(u)), "U-Net transpose");
  HCHK(huvk::launch_conv3x3_wgrad(a.dyt, a.xt, a.N, H, W, c.cin, c.cout, splits, a.part, S(u)),
       "U-Net weight gradient");
  HCHK(huvk::launch_sum_splits(a.part, splits, (int64_t)c.cout * 9 * c.cin, u->g + c.w_off, S(u)),
       "U-Net split sum");
  // bias partials reuse the (already consumed) weight-gradient split buffer
  const int bsplits = bias_splits(c, a.N, H, W);
  HCHK(huvk::launch_channel_sum(a.dyt, a.N, H * W, c.cout, a.part, bsplits, u->g + c.b_off, S(u)),
       "U-Net bias gradient");
  if (dx)
    HCHK(huvk::launch_conv3x3_fwd(dy, a.N, H, W, c.cout, wrot(u, c), nullptr, c.cin, 0, dx, S(u)),
         "U-Net data gradient");
  return HUV_OK;
}

huv_status backward(huv_unet* u) {
  Acts& a = u->a;
  const int L = u->levels;
  const int64_t P = (int64_t)a.N * a.H * a.W;
  HCHK(huvk::launch_head_bwd(a.d2[0], a.dout, P, u->base, u->p + u->head_w, a.gd2[0], a.hpart,
                             kHeadBlocks, S(u)),
       "U-Net head gradient");
  // head dw (base values) and db (1) are contiguous in the parameters
  HCHK(huvk::launch_sum_splits(a.hpart, kHeadBlocks, u->base + 1, u->g + u->head_w, S(u)),
       "U-Net head sum");
  huv_status st;
  for (int l = 0; l <= L - 2; ++l) {
    if ((st = conv_bwd(u, u->dec2[l], a.d1[l], a.d2[l], a.gd2[l], l, a.gd1[l])) != HUV_OK) return st;
    if ((st = conv_bwd(u, u->dec1[l], a.cat[l], a.d1[l], a.gd1[l], l, a.gcat[l])) != HUV_OK) return st;
    bf16_t* gup = l == L - 2 ? a.ge2[L - 1] : a.gd2[l + 1];
    HCHK(huvk::launch_upcat_bwd(a.gcat[l], u->ch(l + 1), u->ch(l), a.N, a.H >> l, a.W >> l, gup,
                                a.ge2[l], 0, S(u)),
         "U-Net upsample gradient");
  }
  for (int l = L - 1; l >= 0; --l) {
    const int H = a.H >> l, W = a.W >> l;
    if (l < L - 1) {
      HCHK(huvk::launch_maxpool_bwd(a.e2[l], a.gpool[l], a.N, H, W, u->ch(l), a.tmp, S(u)),
           "U-Net pool gradient");
      HCHK(huvk::launch_add_bf16(a.ge2[l], a.tmp, (int64_t)a.N * H * W * u->ch(l), S(u)),
           "U-Net skip gradient");
    }
    if ((st = conv_bwd(u, u->enc2[l], a.e1[l], a.e2[l], a.ge2[l], l, a.ge1[l])) != HUV_OK) return st;
    const bf16_t* src = l == 0 ? a.x8 : a.pool[l - 1];
    if ((st = conv_bwd(u, u->enc1[l], src, a.e1[l], a.ge1[l], l, l == 0 ? nullptr : a.gpool[l - 1])) != HUV_OK)
      return st;
  }
  return HUV_OK;
}

huv_status check_shape(huv_unet* u, int N, int H, int W) {
  const int q = 16 << (u->levels - 1);
  if (N < 1 || H < q || W < q || H % q || W % q)
    return fail(HUV_INVALID_PARAMETER, "U-Net input: H and W must be positive multiples of " +
                                           std::to_string(q));
  return HUV_OK;
}

// splitmix64 + Box-Muller normal draws for the initial weights.
struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double normal() {
    const double u1 = uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

}  // namespace

extern "C" {

huv_status huv_unet_create(huv_ctx* c, int levels, int base, uint64_t seed, huv_unet** out) {
  if (!c || !out) return fail(HUV_INVALID_PARAMETER, "null argument");
  if (levels < 2 || levels > 5) return fail(HUV_INVALID_PARAMETER, "levels must be in 2..5");
  if (base != 16 && base != 32 && base != 64) return fail(HUV_INVALID_PARAMETER, "base channels must be 16, 32 or 64");
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  auto u = std::make_unique<huv_unet>();
  u->ctx = c;
  u->levels = levels;
  u->base = base;
  int64_t off = 0;
  auto add = [&](int cin, int cout) {
    Conv k;
    k.cin = cin;
    k.cout = cout;
    k.w_off = off;
    off += (int64_t)cout * 9 * cin;
    k.b_off = off;
    off += cout;
    return k;
  };
  for (int l = 0; l < levels; ++l) {
    u->enc1.push_back(add(l == 0 ? 8 : u->ch(l - 1), u->ch(l)));
    u->enc2.push_back(add(u->ch(l), u->ch(l)));
  }
  u->dec1.resize(levels);
  u->dec2.resize(levels);
  for (int l = levels - 2; l >= 0; --l) {
    u->dec1[l] = add(u->ch(l + 1) + u->ch(l), u->ch(l));
    u->dec2[l] = add(u->ch(l), u->ch(l));
  }
  u->dec1.resize(levels - 1);
  u->dec2.resize(levels - 1);
  u->head_w = off;
  off += base;
  u->head_b = off;
  off += 1;
  u->nparams = off;
  // He initialisation (fan-in 9 cin; the packed input has one live channel).
  std::vector<float> h((size_t)off, 0.f);
  Rng rng{seed ^ 0x5eed0f0e7ull};
  auto init = [&](const Conv& k, bool first) {
    const double sd = std::sqrt(2.0 / (9.0 * (first ? 1 : k.cin)));
    for (int64_t i = 0; i < (int64_t)k.cout * 9 * k.cin; ++i)
      h[k.w_off + i] = (float)(sd * rng.normal());
  };
  for (int l = 0; l < levels; ++l) {
    init(u->enc1[l], l == 0);
    init(u->enc2[l], false);
  }
  for (int l = levels - 2; l >= 0; --l) {
    init(u->dec1[l], false);
    init(u->dec2[l], false);
  }
  for (int i = 0; i < base; ++i) h[u->head_w + i] = (float)(std::sqrt(1.0 / base) * rng.normal());
  const size_t fb = (size_t)off * sizeof(float), bb = (size_t)off * sizeof(bf16_t);
  hipError_t e = hipMalloc(&u->p, fb);
  if (e == hipSuccess) e = hipMalloc(&u->g, fb);
  if (e == hipSuccess) e = hipMalloc(&u->m, fb);
  if (e == hipSuccess) e = hipMalloc(&u->v, fb);
  if (e == hipSuccess) e = hipMalloc(&u->wb, bb);
  if (e == hipSuccess) e = hipMalloc(&u->wr, bb);
  if (e == hipSuccess) e = hipMemcpy(u->p, h.data(), fb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(u->g, 0, fb);
  if (e == hipSuccess) e = hipMemset(u->m, 0, fb);
  if (e == hipSuccess) e = hipMemset(u->v, 0, fb);
  if (e != hipSuccess) return hip_fail(e, "U-Net parameter allocation");
  huv_status st = refresh_bf16(u.get());
  if (st != HUV_OK) return st;
  *out = u.release();
  return HUV_OK;
}

void huv_unet_destroy(huv_unet* u) {
  if (!u) return;
  if (u->ctx) {
    std::lock_guard<std::mutex> lk(u->ctx->mu);
    hipSetDevice(u->ctx->device);
    hipStreamSynchronize(u->ctx->stream());
  }
  delete u;
}

huv_status huv_unet_num_params(const huv_unet* u, int64_t* n) {
  if (!u || !n) return fail(HUV_INVALID_PARAMETER, "null argument");
  *n = u->nparams;
  return HUV_OK;
}

huv_status huv_unet_get_params(huv_unet* u, float* host) {
  if (!u || !host) return fail(HUV_INVALID_PARAMETER, "null argument");
  std::lock_guard<std::mutex> lk(u->ctx->mu);
  hipSetDevice(u->ctx->device);
  HCHK(hipStreamSynchronize(S(u)), "U-Net sync");
  HCHK(hipMemcpy(host, u->p, (size_t)u->nparams * sizeof(float), hipMemcpyDeviceToHost),
       "U-Net parameter copy");
  return HUV_OK;
}

huv_status huv_unet_set_params(huv_unet* u, const float* host) {
  if (!u || !host) return fail(HUV_INVALID_PARAMETER, "null argument");
  std::lock_guard<std::mutex> lk(u->ctx->mu);
  hipSetDevice(u->ctx->device);
  HCHK(hipStreamSynchronize(S(u)), "U-Net sync");
  HCHK(hipMemcpy(u->p, host, (size_t)u->nparams * sizeof(float), hipMemcpyHostToDevice),
       "U-Net parameter copy");
  return refresh_bf16(u);
}

huv_status huv_unet_forward(huv_unet* u, int N, int H, int W, const float* in, float* out) {
  if (!u || !in || !out) return fail(HUV_INVALID_PARAMETER, "null argument");
  huv_status st = check_shape(u, N, H, W);
  if (st != HUV_OK) return st;
  std::lock_guard<std::mutex> lk(u->ctx->mu);
  hipSetDevice(u->ctx->device);
  if ((st = ensure_acts(u, N, H, W)) != HUV_OK) return st;
  if ((st = forward(u, in)) != HUV_OK) return st;
  HCHK(hipMemcpyAsync(out, u->a.out, (size_t)N * H * W * sizeof(float), hipMemcpyDeviceToDevice,
                      S(u)),
       "U-Net output copy");
  return HUV_OK;
}

huv_status huv_unet_loss_grad(huv_unet* u, int N, int H, int W, const float* in,
                              const float* target, double* loss, float* grad_host) {
  if (!u || !in || !target || !loss) return fail(HUV_INVALID_PARAMETER, "null argument");
  huv_status st = check_shape(u, N, H, W);
  if (st != HUV_OK) return st;
  std::lock_guard<std::mutex> lk(u->ctx->mu);
  hipSetDevice(u->ctx->device);
  if ((st = ensure_acts(u, N, H, W)) != HUV_OK) return st;
  if ((st = forward(u, in)) != HUV_OK) return st;
  const int64_t P = (int64_t)N * H * W;
  HCHK(huvk::launch_mse(u->a.out, target, P, u->a.dout, u->a.lpart, kLossBlocks, S(u)), "U-Net loss");
  if ((st = backward(u)) != HUV_OK) return st;
  std::vector<double> lp(kLossBlocks);
  HCHK(hipMemcpyAsync(lp.data(), u->a.lpart, kLossBlocks * sizeof(double), hipMemcpyDeviceToHost, S(u)),
       "U-Net loss copy");
  HCHK(hipStreamSynchronize(S(u)), "U-Net sync");
  double s = 0.0;
  for (double x : lp) s += x;
  *loss = s / (double)P;
  if (grad_host)
    HCHK(hipMemcpy(grad_host, u->g, (size_t)u->nparams * sizeof(float), hipMemcpyDeviceToHost),
         "U-Net gradient copy");
  return HUV_OK;
}

huv_status huv_unet_adam(huv_unet* u, double lr, double beta1, double beta2, double eps) {
  if (!u) return fail(HUV_INVALID_PARAMETER, "null argument");
  if (!(lr > 0.0) || !(beta1 >= 0.0 && beta1 < 1.0) || !(beta2 >= 0.0 && beta2 < 1.0) || !(eps > 0.0))
    return fail(HUV_INVALID_PARAMETER, "Adam: need lr > 0, 0 <= beta < 1, eps > 0");
  std::lock_guard<std::mutex> lk(u->ctx->mu);
  hipSetDevice(u->ctx->device);
  u->t += 1;
  const float c1 = (float)(1.0 - std::pow(beta1, (double)u->t));
  const float c2 = (float)(1.0 - std::pow(beta2, (double)u->t));
  HCHK(huvk::launch_adam(u->p, u->g, u->m, u->v, u->nparams, (float)lr, (float)beta1,
                         (float)beta2, (float)eps, c1, c2, S(u)),
       "U-Net Adam");
  return refresh_bf16(u);
}

}  // extern "C"
