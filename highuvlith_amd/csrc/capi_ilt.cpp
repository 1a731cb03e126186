// C ABI of the GPU ILT (include/huv.h): batched cost / exact adjoint gradient
// (IltProblem::cost_and_gradient_with, ilt.rs:667-738) and the CG / Armijo
// driver (run_stage, ilt.rs:755-866) with per-target state on the host and
// all vector work on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/huv.h"
#include "capi_internal.hpp"
#include "hip/ilt_kernels.hpp"
#include "hip/socs_kernels.hpp"

using namespace huvi;

struct huv_ilt {
  huv_ctx* ctx = nullptr;
  int n = 0, ncond = 0, B = 0, m = 0;
  std::vector<huv_kset*> conds;
  huv_ilt_config cfg{};
  double flare = 0.0;
  bool complex_abs = false;  // absorber with an imaginary part: complex transmittance
  bool proxy = false;        // local-proxy gradient
  int Bf = 0;                // density-filter band half-width (0: no filter)
  float filter_c = 0.0f;     // filter exponent per squared bin index, 2 pi^2 s^2 df^2
  float* dose_d = nullptr;
  float* weight_d = nullptr;
  std::vector<float> weight_h;  // host copy of the condition weights
  ~huv_ilt() {
    if (dose_d) hipFree(dose_d);
    if (weight_d) hipFree(weight_d);
  }
};

namespace {

constexpr int kBpi = 64;  // reduction blocks per image

struct IltBuffers {
  float *m, *t, *G, *dcdm, *reg_grad, *G0g;
  float2 *Gb, *Rout, *H;
  float2* tc = nullptr;                       // complex transmittance (complex absorber)
  float *thf = nullptr, *gm = nullptr;        // filtered theta, chain-rule temporary (filter)
  float2 *Xf = nullptr, *If = nullptr;        // filter band box and half band
  float2 *g0tmp = nullptr, *abuf = nullptr;   // adjoint on coarse grids above 64
  double* proxy_part = nullptr;               // per-image sums of G (local proxy)
  double *fid_part, *reg_part, *cost_d;
  int KS = 1;  // adjoint work-groups per image (kernels dealt over them)
};

// One adjoint work-group per kernel of the largest condition.
static int ilt_splits(const huv_ilt* p) {
  int ks = 1;
  for (const huv_kset* k : p->conds) ks = std::max(ks, k->K);
  return ks;
}

huv_status ilt_buffers(Ctx* c, const huv_ilt* p, int T, IltBuffers& b) {
  hipError_t e;
  const size_t N = (size_t)p->n * p->n;
  const int W = 2 * p->B + 1, W2 = 4 * p->B + 1;
  const size_t TC = (size_t)T * p->ncond;
  // m, then t (real) or tc (interleaved complex)
  const size_t tw = p->complex_abs ? 2 : 1;
  char* a = (char*)c->ws.get(Workspace::ILT0, (1 + tw) * (size_t)T * N * sizeof(float), e);
  if (!a) return hip_fail(e, "ILT workspace");
  b.m = (float*)a;
  b.t = b.m + (size_t)T * N;
  b.tc = p->complex_abs ? (float2*)b.t : nullptr;
  b.G = (float*)c->ws.get(Workspace::ILT2, TC * N * sizeof(float), e);
  if (!b.G) return hip_fail(e, "ILT workspace");
  a = (char*)c->ws.get(Workspace::ILT3, 2 * (size_t)T * N * sizeof(float), e);
  if (!a) return hip_fail(e, "ILT workspace");
  b.dcdm = (float*)a;
  b.reg_grad = b.dcdm + (size_t)T * N;
  b.KS = ilt_splits(p);
  const size_t nGb = TC * W2 * W2, nR = TC * b.KS * W * W, nH = (size_t)T * (p->B + 1) * W;
  const size_t nG0 = TC * p->m * p->m;  // floats
  a = (char*)c->ws.get(Workspace::ILT4, (nGb + nR + nH) * sizeof(float2) + nG0 * sizeof(float), e);
  if (!a) return hip_fail(e, "ILT workspace");
  b.Gb = (float2*)a;
  b.Rout = b.Gb + nGb;
  b.H = b.Rout + nR;
  b.G0g = (float*)(b.H + nH);
  a = (char*)c->ws.get(Workspace::ILT5, (TC * kBpi + (size_t)T * kBpi + T) * sizeof(double), e);
  if (!a) return hip_fail(e, "ILT workspace");
  b.fid_part = (double*)a;
  b.reg_part = b.fid_part + TC * kBpi;
  b.cost_d = b.reg_part + (size_t)T * kBpi;
  if (p->m > 64) {
    const size_t ng0 = TC * W2 * p->m, nab = TC * b.KS * W * p->m;
    a = (char*)c->ws.get(Workspace::ILT9, (ng0 + nab) * sizeof(float2), e);
    if (!a) return hip_fail(e, "ILT workspace");
    b.g0tmp = (float2*)a;
    b.abuf = b.g0tmp + ng0;
  }
  if (p->Bf > 0) {
    const int Wf = 2 * p->Bf + 1;
    const size_t nx = (size_t)T * Wf * Wf, ni = (size_t)T * (p->Bf + 1) * Wf;
    a = (char*)c->ws.get(Workspace::ILT10,
                         2 * (size_t)T * N * sizeof(float) + (nx + ni) * sizeof(float2), e);
    if (!a) return hip_fail(e, "ILT workspace");
    b.thf = (float*)a;
    b.gm = b.thf + (size_t)T * N;
    b.Xf = (float2*)(b.gm + (size_t)T * N);
    b.If = b.Xf + nx;
  }
  if (p->proxy) {
    b.proxy_part = (double*)c->ws.get(Workspace::ILT11, TC * kBpi * sizeof(double), e);
    if (!b.proxy_part) return hip_fail(e, "ILT workspace");
  }
  return HUV_OK;
}

// out = G_s * in for count real n x n fields (the density filter, applied on
// the |k| <= Bf band where its spectrum exceeds 1e-12; ilt.rs:524-535).
huv_status ilt_filter(huv_ilt* p, int count, const float* in, float* out, IltBuffers& b) {
  Ctx* c = p->ctx;
  const int n = p->n;
  huv_status st = band_stage(c, count, n, p->Bf, p->Bf, in, b.Xf, false);
  if (st != HUV_OK) return st;
  hipError_t e = huvk::launch_ilt_filter_band(b.Xf, p->Bf, p->filter_c,
                                              (float)(1.0 / ((double)n * n)), count, b.If,
                                              c->stream());
  if (e != hipSuccess) return hip_fail(e, "ilt filter band launch");
  const float2* tw_n = c->twiddles(n, e);
  if (!tw_n) return hip_fail(e, "twiddle table");
  e = huvk::launch_band_image(n, p->Bf, count, b.If, tw_n, out, c->stream());
  return e == hipSuccess ? HUV_OK : hip_fail(e, "ilt filter image launch");
}

// Where the theta rows of an evaluation come from: row j is
// theta[rows[j]] + coef[j] dir[rows[j]] (rows null: row j; coef null: theta
// alone), read in place.  stage: room for count compact rows, written only
// when the density filter needs the rows as fields.
struct ThetaRows {
  const float* theta = nullptr;
  const int* rows = nullptr;
  const float* dir = nullptr;
  const double* coef = nullptr;
  float* stage = nullptr;
};

// m (may be null) and the transmittance t (may be null; complex with a
// complex absorber) of count evaluation rows: the density filter first when
// configured.
huv_status ilt_mask_t(huv_ilt* p, int count, const ThetaRows& src, float* m, void* t,
                      IltBuffers& b) {
  Ctx* c = p->ctx;
  hipStream_t s = c->stream();
  const int64_t N = (int64_t)p->n * p->n;
  const float beta = (float)p->cfg.beta, ar = (float)p->cfg.absorber_re;
  const float ai = (float)p->cfg.absorber_im;
  float* tr = p->complex_abs ? nullptr : (float*)t;
  float2* tc = p->complex_abs ? (float2*)t : nullptr;
  hipError_t e;
  if (p->Bf > 0) {
    // the filter acts on theta as a field: materialise the rows first
    const float* th = src.theta;
    if (src.coef || src.rows) {
      if (!src.stage || (src.coef && !src.rows))
        return fail(HUV_INVALID_PARAMETER, "filtered evaluation rows need a staging buffer");
      e = src.coef ? huvk::launch_vec_axpy_gather(src.theta, src.dir, src.coef, N, count, src.rows,
                                                  src.stage, s)
                   : huvk::launch_rows_gather(src.theta, src.rows, N, count, src.stage, s);
      if (e != hipSuccess) return hip_fail(e, "ilt theta rows");
      th = src.stage;
    }
    huv_status st = ilt_filter(p, count, th, b.thf, b);
    if (st != HUV_OK) return st;
    e = huvk::launch_ilt_mask_rows(b.thf, nullptr, nullptr, nullptr, N, count, beta, ar, ai, m, tr,
                                   tc, s);
  } else {
    e = huvk::launch_ilt_mask_rows(src.theta, src.rows, src.dir, src.coef, N, count, beta, ar, ai,
                                   m, tr, tc, s);
  }
  return e == hipSuccess ? HUV_OK : hip_fail(e, "ilt mask launch");
}

// Cost of T evaluation rows (into b.cost_d, device; device_cost false: only
// the per-row partial sums b.fid_part and b.reg_part, for ilt_host_costs) and
// optionally the theta-gradient.  Row j compares with target row trows[j]
// (null: j) and its gradient goes to grad row grows[j] (null: j).
huv_status ilt_eval(huv_ilt* p, int T, const float* targets, const int* trows,
                    const ThetaRows& src, float* grad, const int* grows, IltBuffers& b,
                    bool device_cost = true) {
  Ctx* c = p->ctx;
  hipStream_t s = c->stream();
  hipError_t e;
  const int n = p->n, C = p->ncond, B = p->B, m = p->m;
  const int64_t N = (int64_t)n * n;
  const bool reg = p->cfg.tv_weight > 0.0 || p->cfg.bin_weight > 0.0;
  hipEvent_t ev;
  c->stage_begin(ST_ILT_POINTWISE, &ev);
  huv_status st = ilt_mask_t(p, T, src, grad || reg ? b.m : nullptr,
                             p->complex_abs ? (void*)b.tc : (void*)b.t, b);
  if (st != HUV_OK) return st;
  c->stage_end(ST_ILT_POINTWISE, ev);
  float2* Mq = nullptr;
  MaskSource mask = MaskSource::real(b.t);
  if (p->complex_abs) mask.kind = MaskSource::COMPLEX;
  // The kernel sum is split as for a single target at every batch size, so a
  // target's images, and with them its whole optimisation trajectory, do not
  // depend on which other targets are still live in the batch.  The images
  // are never stored: the last stage evaluates the fidelity (and G) of every
  // pixel where it would have written it, with per-row-class partial sums.
  huvk::FidelityEpi fid;
  fid.target = targets;
  fid.trows = trows;
  fid.ncond = C;
  fid.cost_kind = p->cfg.cost_kind;
  fid.threshold = (float)p->cfg.threshold;
  fid.steepness = (float)p->cfg.steepness;
  fid.inv_n = 1.0f / (float)N;
  fid.dose = p->dose_d;
  fid.G = grad ? b.G : nullptr;
  fid.partial = b.fid_part;
  st = socs_forward(c, T, n, mask, C, p->conds.data(), nullptr, false, &Mq, nullptr, 0, 1, &fid);
  if (st != HUV_OK) return st;
  c->stage_begin(ST_ILT_POINTWISE, &ev);
  if (reg) {
    e = huvk::launch_ilt_reg(b.m, n, T, kBpi, (float)p->cfg.tv_weight, (float)p->cfg.tv_epsilon,
                             (float)p->cfg.bin_weight, b.reg_grad, b.reg_part, s);
    if (e != hipSuccess) return hip_fail(e, "ilt reg launch");
  }
  if (device_cost) {
    e = huvk::launch_ilt_cost_reduce(b.fid_part, reg ? b.reg_part : nullptr, T, C, n / 32, kBpi,
                                     p->weight_d, 1.0 / (double)N, b.cost_d, s);
    if (e != hipSuccess) return hip_fail(e, "ilt cost reduce launch");
  }
  c->stage_end(ST_ILT_POINTWISE, ev);
  if (!grad) return HUV_OK;
  const bool filt = p->Bf > 0;
  bool chained = false;  // the chain rule ran in the gradient image stage
  if (p->proxy) {
    // local proxy: dC/dm = sum_c w_c G0_c (ilt.rs:719-721)
    c->stage_begin(ST_ILT_OTHER, &ev);
    e = huvk::launch_ilt_proxy(b.G, T, C, N, kBpi, p->weight_d, (float)p->flare, b.proxy_part,
                               b.dcdm, s);
    if (e != hipSuccess) return hip_fail(e, "ilt proxy launch");
    c->stage_end(ST_ILT_OTHER, ev);
  } else {
    // |g| <= 2B band of F(G) for every (target, condition) image.
    st = band_stage(c, T * C, n, 2 * B, 2 * B, b.G, b.Gb, false);
    if (st != HUV_OK) return st;
    const int* dK;
    const float2* const* dkern;
    const float* const* dlam;
    st = plane_table(c, C, p->conds.data(), &dK, &dkern, &dlam);
    if (st != HUV_OK) return st;
    c->stage_begin(ST_ILT_ADJOINT, &ev);
    const double n2 = (double)n * n;
    const float g0_scale = 1.0f / ((float)m * (float)m);
    const float lam_scale = (float)n2;  // device lambda carries 1/n^4
    if (m <= 64) {
      e = huvk::launch_ilt_adjoint(m, T * C, b.KS, Mq, b.Gb, B, C, dK, dkern, dlam, lam_scale,
                                   g0_scale, (float)p->flare, b.G0g, b.Rout, s);
    } else {
      if ((int64_t)T * C * b.KS > 65535)
        return fail(HUV_INVALID_PARAMETER, "targets x conditions x kernels exceeds 65535 per "
                                           "evaluation on coarse grids above 64: pass fewer targets");
      const float2* tw_m = c->twiddles(m, e);
      if (!tw_m) return hip_fail(e, "twiddle table");
      e = huvk::launch_ilt_adjoint_big(m, T * C, b.KS, Mq, b.Gb, B, C, dK, dkern, dlam,
                                       lam_scale, g0_scale, (float)p->flare, tw_m, b.g0tmp,
                                       b.G0g, b.abuf, b.Rout, s);
    }
    if (e != hipSuccess) return hip_fail(e, "ilt adjoint launch");
    c->stage_end(ST_ILT_ADJOINT, ev);
    c->stage_begin(ST_ILT_OTHER, &ev);
    e = huvk::launch_ilt_hband(b.Rout, B, T, C, b.KS, p->weight_d,
                               (float)(1.0 - p->cfg.absorber_re), (float)(-p->cfg.absorber_im),
                               (float)(1.0 / n2), b.H, s);
    if (e != hipSuccess) return hip_fail(e, "ilt hband launch");
    const float2* tw_n = c->twiddles(n, e);
    if (!tw_n) return hip_fail(e, "twiddle table");
    if (filt) {
      e = huvk::launch_band_image(n, B, T, b.H, tw_n, b.dcdm, s);
    } else {
      // the chain rule where the image of dC/dm would be stored: the
      // gradient rows are written directly
      huvk::ChainEpi ch;
      ch.m = b.m;
      ch.reg = reg ? b.reg_grad : nullptr;
      ch.orows = grows;
      ch.beta = (float)p->cfg.beta;
      ch.grad = grad;
      e = huvk::launch_band_image_chain(n, B, T, b.H, tw_n, ch, s);
      chained = true;
    }
    if (e != hipSuccess) return hip_fail(e, "ilt band image launch");
    c->stage_end(ST_ILT_OTHER, ev);
  }
  // chain through the sigmoid, then the (self-adjoint) density filter
  c->stage_begin(ST_ILT_OTHER, &ev);
  if (!chained) {
    e = huvk::launch_ilt_chain(b.dcdm, reg ? b.reg_grad : nullptr, b.m, (float)p->cfg.beta, N, T,
                               filt ? nullptr : grows, filt ? b.gm : grad, s);
    if (e != hipSuccess) return hip_fail(e, "ilt chain launch");
  }
  if (filt) {
    // the filter writes compact rows; those of a row subset go out through thf
    st = ilt_filter(p, T, b.gm, grows ? b.thf : grad, b);
    if (st != HUV_OK) return st;
    if (grows) {
      e = huvk::launch_rows_scatter(b.thf, grows, N, T, grad, s);
      if (e != hipSuccess) return hip_fail(e, "ilt gradient scatter");
    }
  }
  c->stage_end(ST_ILT_OTHER, ev);
  return HUV_OK;
}

// Costs of `count` evaluation rows from their partial sums (fid: count x ncond
// x bpi, reg: count x kBpi or null), with the arithmetic of
// k_ilt_cost_reduce: per condition the partials in order, weighted and scaled
// with one fused multiply-add, then the regularisation partials in order.
void ilt_host_costs(const huv_ilt* p, int count, const double* fid, const double* reg,
                    double* cost) {
  const int C = p->ncond, bpi = p->n / 32;
  const double inv_n = 1.0 / ((double)p->n * p->n);
  for (int t = 0; t < count; ++t) {
    double total = 0.0;
    for (int c = 0; c < C; ++c) {
      const double* q = fid + ((size_t)t * C + c) * bpi;
      double s = 0.0;
      for (int b = 0; b < bpi; ++b) s += q[b];
      total = std::fma((double)p->weight_h[c] * s, inv_n, total);
    }
    if (reg)
      for (int b = 0; b < kBpi; ++b) total += reg[(size_t)t * kBpi + b];
    cost[t] = total;
  }
}

}  // namespace

extern "C" {

huv_status huv_ilt_create(huv_ctx* c, int n, int ncond, huv_kset* const* conds, const double* dose,
                          const double* weight, const huv_ilt_config* cfg, huv_ilt** out) {
  if (!c || !conds || !dose || !weight || !cfg || !out)
    return fail(HUV_INVALID_PARAMETER, "null argument");
  if (ncond < 1) return fail(HUV_INVALID_PARAMETER, "conditions: at least one process condition is required");
  huv_status st = validate_planes(c, n, ncond, conds);
  if (st != HUV_OK) return st;
  double wsum = 0.0;
  for (int i = 0; i < ncond; ++i) {
    if (!(std::isfinite(dose[i]) && dose[i] > 0.0)) return fail(HUV_INVALID_PARAMETER, "dose: must be positive and finite");
    if (!(std::isfinite(weight[i]) && weight[i] >= 0.0)) return fail(HUV_INVALID_PARAMETER, "weight: must be non-negative and finite");
    wsum += weight[i];
  }
  if (wsum <= 0.0) return fail(HUV_INVALID_PARAMETER, "weight: condition weights must not all be zero");
  if (!(cfg->beta > 0.0 && std::isfinite(cfg->beta))) return fail(HUV_INVALID_PARAMETER, "sigmoid_steepness: must be positive and finite");
  if (!(cfg->tv_epsilon > 0.0)) return fail(HUV_INVALID_PARAMETER, "tv_epsilon: must be positive and finite");
  if (!(cfg->tv_weight >= 0.0 && cfg->bin_weight >= 0.0)) return fail(HUV_INVALID_PARAMETER, "regularization weights must be non-negative");
  if (cfg->cost_kind == 1 && !(cfg->threshold > 0.0 && cfg->steepness > 0.0))
    return fail(HUV_INVALID_PARAMETER, "threshold / steepness: must be positive and finite");
  if (!(std::isfinite(cfg->absorber_re) && std::isfinite(cfg->absorber_im)))
    return fail(HUV_INVALID_PARAMETER, "absorber: must be finite");
  if (!(cfg->min_feature_nm >= 0.0 && std::isfinite(cfg->min_feature_nm)))
    return fail(HUV_INVALID_PARAMETER, "min_feature_nm: must be non-negative and finite");
  if (cfg->gradient != 0 && cfg->gradient != 1)
    return fail(HUV_INVALID_PARAMETER, "gradient: 0 (adjoint) or 1 (local proxy)");
  if (n < 256) return fail(HUV_UNSUPPORTED, "GPU ILT needs grids of 256 or more");
  // The adjoint works on the stored kernel boxes (half-width coarse_band(m)).
  const int m = conds[0]->m, B = conds[0]->BB;
  if (m > 256 || 4 * B + 1 > n)
    return fail(HUV_UNSUPPORTED, "GPU ILT adjoint supports coarse grids up to 256 (band <= 63)");
  // Density filter (ilt.rs:443-454): exp(-2 pi^2 s^2 f^2), s = min_feature / 1.349,
  // kept on the bins where it exceeds 1e-12.
  int Bf = 0;
  float filter_c = 0.0f;
  if (cfg->min_feature_nm > 0.0) {
    if (!(cfg->pixel_nm > 0.0 && std::isfinite(cfg->pixel_nm)))
      return fail(HUV_INVALID_PARAMETER, "pixel_nm: must be positive with a density filter");
    const double s = cfg->min_feature_nm / 1.349, df = 1.0 / (n * cfg->pixel_nm);
    const double cc = 2.0 * M_PI * M_PI * s * s * df * df;
    const double kmax = std::sqrt(std::log(1e12) / cc);
    if (kmax > 127.0 || 2.0 * std::floor(kmax) + 1.0 > n)
      return fail(HUV_UNSUPPORTED, "density filter wider than the 127-bin band (min_feature_nm "
                                   "too small for this grid)");
    Bf = std::max(1, (int)std::floor(kmax));
    filter_c = (float)cc;
  }
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  auto p = std::make_unique<huv_ilt>();
  p->ctx = c;
  p->n = n;
  p->ncond = ncond;
  p->B = B;
  p->m = m;
  p->conds.assign(conds, conds + ncond);
  p->cfg = *cfg;
  p->flare = conds[0]->flare;
  p->complex_abs = cfg->absorber_im != 0.0;
  p->proxy = cfg->gradient == 1;
  p->Bf = Bf;
  p->filter_c = filter_c;
  std::vector<float> d(dose, dose + ncond), w(weight, weight + ncond);
  p->weight_h = w;
  hipError_t e = hipMalloc(&p->dose_d, ncond * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&p->weight_d, ncond * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(p->dose_d, d.data(), ncond * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->weight_d, w.data(), ncond * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "ILT condition tables");
  *out = p.release();
  return HUV_OK;
}

void huv_ilt_destroy(huv_ilt* p) { delete p; }

huv_status huv_ilt_cost_grad(huv_ilt* p, int T, const float* targets, const float* theta,
                             double* cost, float* grad) {
  if (!p || !targets || !theta || !cost) return fail(HUV_INVALID_PARAMETER, "null argument");
  if (T < 1) return fail(HUV_INVALID_PARAMETER, "need at least one target");
  Ctx* c = p->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  IltBuffers b;
  huv_status st = ilt_buffers(c, p, T, b);
  if (st != HUV_OK) return st;
  ThetaRows src;
  src.theta = theta;
  st = ilt_eval(p, T, targets, nullptr, src, grad, nullptr, b);
  if (st != HUV_OK) return st;
  hipError_t e = hipMemcpyAsync(cost, b.cost_d, T * sizeof(double), hipMemcpyDeviceToHost, c->stream());
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream());
  return e == hipSuccess ? HUV_OK : hip_fail(e, "cost copy");
}

}  // extern "C"

namespace {

// Optional trace of a stage: gradient norms and mask snapshots.
struct StageTrace {
  double* gnorm_hist = nullptr;
  int every = 0;
  const int* offset = nullptr;
  int nslots = 0;
  float* snapshots = nullptr;
};

// run_stage for every target; stop_cost (nullable) ends a target as soon as
// its cost is at or below stop_cost[t] (termination 3).
huv_status ilt_optimize(huv_ilt* p, int T, const float* targets, float* theta, int max_iterations,
                        double learning_rate, double tol, int use_cg, const double* stop_cost,
                        double* cost_hist, int* iterations, int* termination,
                        const StageTrace& tr = StageTrace()) {
  if (!p || !targets || !theta || !cost_hist || !iterations || !termination)
    return fail(HUV_INVALID_PARAMETER, "null argument");
  if (T < 1 || max_iterations < 0) return fail(HUV_INVALID_PARAMETER, "bad target count or budget");
  if (!(learning_rate > 0.0)) return fail(HUV_INVALID_PARAMETER, "learning_rate: must be positive and finite");
  if (tr.every < 0 || (tr.every > 0 && (tr.nslots < 1 || !tr.snapshots)))
    return fail(HUV_INVALID_PARAMETER, "snapshots: need a buffer of at least one slot");
  Ctx* c = p->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  hipStream_t s = c->stream();
  IltBuffers b;
  huv_status st = ilt_buffers(c, p, T, b);
  if (st != HUV_OK) return st;
  const int64_t N = (int64_t)p->n * p->n;
  const size_t vb = (size_t)T * N * sizeof(float);
  hipError_t e;
  char* vec = (char*)c->ws.get(Workspace::ILT6, 3 * vb, e);
  if (!vec) return hip_fail(e, "ILT optimizer workspace");
  float* grad = (float*)vec;
  float* gnew = (float*)(vec + vb);
  float* dir = (float*)(vec + 2 * vb);
  const int nblk = 64;
  constexpr int kRed = 3;  // reductions fetched together

  auto fail_e = [&](const char* what) { return hip_fail(e, what); };
  std::vector<double> h_red((size_t)kRed * T), h_a(T), h_b(T);
  // Small uploads (coefficients, row lists) are appended to the pinned
  // staging area and copied with one transfer before the launches that read
  // them; the area is reused after every stream synchronisation.
  const size_t stage_cap = 96 * (size_t)T + 4096;
  if (!c->staging(stage_cap, e)) return fail_e("staging buffer");
  // Reduction and cost partials are written by the kernels straight into
  // pinned, device-mapped host memory (no copy launch): the host reads them
  // after the synchronisation that ends the launch.  One area serves both;
  // every use is read before the next launch that writes it.
  const bool has_reg = p->cfg.tv_weight > 0.0 || p->cfg.bin_weight > 0.0;
  const size_t fid_n = (size_t)p->ncond * (p->n / 32);  // fidelity partials per row
  // Evaluation partials first, reduction partials after them, so an
  // evaluation and a reduction can share one synchronisation.
  const size_t red_off = ((size_t)T * (fid_n + kBpi) + 31) & ~(size_t)31;
  const size_t dl_count = red_off + (size_t)kRed * T * nblk;
  if (!c->download_area(dl_count * sizeof(double), e)) return fail_e("download buffer");
  double* dl = (double*)c->dl_h;
  double* dl_dev = (double*)c->dl_d;
  double* dlr = dl + red_off;          // reduction partials (host view)
  double* dlr_dev = dl_dev + red_off;  // the same, device view
  b.fid_part = dl_dev;
  b.reg_part = dl_dev + (size_t)T * fid_n;
  size_t so = 0, sf = 0;
  auto flush = [&]() -> bool {
    if (so > sf) {
      e = hipMemcpyAsync(c->pin_d + sf, c->pin_h + sf, so - sf, hipMemcpyHostToDevice, s);
      if (e != hipSuccess) return false;
      sf = so;
    }
    return true;
  };
  auto synced = [&]() { so = sf = 0; };  // after a stream synchronisation
  // costs of `count` evaluation rows (ilt_eval with device_cost false) from
  // their partial sums, after one synchronisation
  // Rows whose max |d| and d.g partials a fused direction update left in the
  // download area (valid until the area is written again).
  std::vector<char> pre_row(T, 0);
  bool pre_valid = false;
  // host_costs reads what the last evaluation left, after a synchronisation
  auto host_costs = [&](int count, double* dst) {
    ilt_host_costs(p, count, dl, has_reg ? dl + (size_t)T * fid_n : nullptr, dst);
  };
  auto get_costs = [&](int count, double* dst) -> bool {
    e = hipStreamSynchronize(s);
    synced();
    if (e != hipSuccess) return false;
    host_costs(count, dst);
    return true;
  };
  // One synchronisation epoch stages at most ~32 T bytes plus alignment.
  auto put = [&](const void* src, size_t bytes) -> void* {
    const size_t a = (so + 15) & ~(size_t)15;
    if (a + bytes > stage_cap) {
      e = hipErrorOutOfMemory;
      return nullptr;
    }
    std::memcpy(c->pin_h + a, src, bytes);
    so = a + bytes;
    return c->pin_d + a;
  };
  // Vector operations touch only the listed target rows (null: all rows).
  auto stage_rows = [&](const std::vector<int>& rows) -> const int* {
    if ((int)rows.size() == T || rows.empty()) return nullptr;
    return (const int*)put(rows.data(), rows.size() * sizeof(int));
  };
  // Up to kRed reductions of two vectors a, b over `rows` in one pass (op codes of
  // launch_vec_reduce: 0 a.a, 1 a.b, 2 b.b, 3 max |a|), one synchronisation;
  // h_red[r * T + t] holds reduction r for each listed t: the device writes
  // per-block partials and the host adds them in block order (the order the
  // device's own finishing kernel uses), which saves a launch per reduction.
  enum { AA = 0, AB = 1, BB = 2, MAXA = 3 };
  // h_red from the partials in the download area, after one synchronisation
  auto sum_partials = [&](std::initializer_list<int> ops, const std::vector<int>& rows) -> bool {
    e = hipStreamSynchronize(s);
    synced();
    if (e != hipSuccess) return false;
    int q = 0;
    for (int op : ops) {
      for (int t : rows) {
        const double* pk = dlr + ((size_t)q * T + t) * nblk;
        double acc = 0.0;
        for (int k = 0; k < nblk; ++k) acc = op == MAXA ? std::fmax(acc, pk[k]) : acc + pk[k];
        h_red[(size_t)q * T + t] = acc;
      }
      ++q;
    }
    return true;
  };
  auto reduce = [&](const float* a, const float* bv, std::initializer_list<int> ops,
                    const std::vector<int>& rows) -> bool {
    if (rows.empty()) return true;
    pre_valid = false;
    const int* rd = stage_rows(rows);
    if (e != hipSuccess || !flush()) return false;
    int code = 0, r = 0;
    for (int op : ops) code |= op << (4 * r++);
    e = huvk::launch_vec_reduce(a, bv, N, (int)rows.size(), rd, nblk, code, r, T, dlr_dev,
                                nullptr, s);
    if (e != hipSuccess) return false;
    return sum_partials(ops, rows);
  };
  // d = a g + b d on the listed rows (a, b from h_a, h_b)
  auto dir_rows = [&](const float* g, const std::vector<int>& rows) -> bool {
    if (rows.empty()) return true;
    const double* ca = (const double*)put(h_a.data(), T * sizeof(double));
    const double* cb = (const double*)put(h_b.data(), T * sizeof(double));
    const int* rd = stage_rows(rows);
    if (e != hipSuccess || !ca || !cb || !flush()) return false;
    e = huvk::launch_vec_dir(g, ca, cb, N, (int)rows.size(), rd, dir, s);
    return e == hipSuccess;
  };
  // dir_rows followed by what the next iteration first asks of the new d,
  // max |d| and d.g (reduce(dir, g, {MAXA, AB})), computed in the same pass;
  // dir_reduce_take then only sums the partials.
  auto dir_reduce_rows = [&](const float* g, const std::vector<int>& rows) -> bool {
    pre_valid = false;
    if (rows.empty()) return true;
    const double* ca = (const double*)put(h_a.data(), T * sizeof(double));
    const double* cb = (const double*)put(h_b.data(), T * sizeof(double));
    const int* rd = stage_rows(rows);
    if (e != hipSuccess || !ca || !cb || !flush()) return false;
    e = huvk::launch_vec_dir_reduce(g, ca, cb, N, (int)rows.size(), rd, nblk, T, dir, dlr_dev, s);
    if (e != hipSuccess) return false;
    std::fill(pre_row.begin(), pre_row.end(), 0);
    for (int t : rows) pre_row[t] = 1;
    pre_valid = true;
    return true;
  };
  // reduce(dir, grad, {MAXA, AB}, rows), from the fused update's partials
  // when it covered every row
  auto dir_reduce_take = [&](const std::vector<int>& rows) -> bool {
    bool have = pre_valid;
    for (int t : rows) have = have && pre_row[t];
    if (rows.empty() || !have) return reduce(dir, grad, {MAXA, AB}, rows);
    pre_valid = false;
    return sum_partials({MAXA, AB}, rows);
  };
  // Targets that are done cost nothing: evaluations read the rows of the
  // targets in `which` in place through a row list and write their gradients
  // back the same way.  th_c: compact theta rows for the density filter;
  // g_c: compact masks for snapshots.
  char* cmp = (char*)c->ws.get(Workspace::ILT8, 2 * vb, e);
  if (!cmp) return hip_fail(e, "ILT optimizer workspace");
  float* th_c = (float*)cmp;
  float* g_c = (float*)(cmp + vb);
  std::vector<double> cost_c(T);
  // defer: launch only; take_costs fills `out` after a later synchronisation
  auto eval_sel = [&](const float* th, float* g, std::vector<double>& out,
                      const std::vector<int>& which, bool defer = false) -> huv_status {
    const int Tc = (int)which.size();
    if (Tc == 0) return HUV_OK;
    if (Tc == T) {
      if (!flush()) return hip_fail(e, "staging copy");
      ThetaRows src;
      src.theta = th;
      huv_status r = ilt_eval(p, T, targets, nullptr, src, g, nullptr, b, false);
      if (r != HUV_OK || defer) return r;
      return get_costs(T, out.data()) ? HUV_OK : hip_fail(e, "cost copy");
    }
    const int* idx = stage_rows(which);
    if (!idx || !flush()) return hip_fail(e, "row staging");
    ThetaRows src;
    src.theta = th;
    src.rows = idx;
    src.stage = th_c;
    huv_status r = ilt_eval(p, Tc, targets, idx, src, g, idx, b, false);
    if (r != HUV_OK || defer) return r;
    if (!get_costs(Tc, cost_c.data())) return hip_fail(e, "cost copy");
    for (int j = 0; j < Tc; ++j) out[which[j]] = cost_c[j];
    return HUV_OK;
  };
  auto take_costs = [&](std::vector<double>& out, const std::vector<int>& which) {
    const int Tc = (int)which.size();
    if (Tc == 0) return;
    if (Tc == T) {
      host_costs(T, out.data());
      return;
    }
    host_costs(Tc, cost_c.data());
    for (int j = 0; j < Tc; ++j) out[which[j]] = cost_c[j];
  };
  std::vector<int> all_t(T);
  for (int t = 0; t < T; ++t) all_t[t] = t;
  auto eval = [&](const float* th, float* g, std::vector<double>& out) -> huv_status {
    return eval_sel(th, g, out, all_t);
  };
  std::vector<int> which;
  // line-search trial rows: target, coefficient, step; per-target first row and count
  std::vector<int> prow, pfirst(T), pcount(T), used(T), acc_dst;
  std::vector<double> acc_coef;
  std::vector<double> pcoef, pstep;

  const int H = max_iterations + 1;
  std::vector<double> cost(T), cnew(T), step(T, learning_rate), slope(T), dmax(T),
      s_try(T), s_used(T), gnorm(T);
  std::vector<int> active(T, 1), small(T, 0), it(T, 0), term(T, 1), attempt(T, 0), searching(T, 0),
      accepted(T, 0);
  for (int t = 0; t < T; ++t)
    for (int k = 0; k < H; ++k) {
      cost_hist[(size_t)t * H + k] = NAN;
      if (tr.gnorm_hist) tr.gnorm_hist[(size_t)t * H + k] = NAN;
    }
  st = eval(theta, grad, cost);
  if (st != HUV_OK) return st;
  // per-target budget: the iterations left after earlier stages
  std::vector<int> limit(T, max_iterations);
  for (int t = 0; t < T; ++t) {
    if (tr.offset) limit[t] = max_iterations - tr.offset[t];
    cost_hist[(size_t)t * H] = cost[t];
    if (limit[t] <= 0) active[t] = 0;  // budget spent: MaxIterations
    if (stop_cost && cost[t] <= stop_cost[t]) {
      active[t] = 0;
      term[t] = 3;
    }
  }
  if (!reduce(grad, nullptr, {AA}, all_t)) return fail_e("dot");
  // g.g of each target's current gradient: the previous iteration's g'.g'
  // (same row, same kernel and order), so the CG update reduces only g.g'
  // and g'.g'
  std::vector<double> gsq(h_red.begin(), h_red.begin() + T);
  for (int t = 0; t < T; ++t) {
    gnorm[t] = std::sqrt(h_red[t]);
    if (tr.gnorm_hist) tr.gnorm_hist[(size_t)t * H] = gnorm[t];
  }
  for (int t = 0; t < T; ++t) { h_a[t] = -1.0; h_b[t] = 0.0; }
  if (!dir_reduce_rows(grad, all_t)) return fail_e("dir");
  // the local proxy runs steepest descent and accepts on plain decrease (ilt.rs:767-801)
  const bool proxy = p->proxy;
  const bool cg = use_cg != 0 && !proxy;
  // mask snapshots: global iteration of target t, and the rows due after this iteration
  auto slot_of = [&](int t) {
    const int g = (tr.offset ? tr.offset[t] : 0) + it[t];
    return (tr.every > 0 && g > 0 && g % tr.every == 0 && g / tr.every - 1 < tr.nslots)
               ? g / tr.every - 1 : -1;
  };
  std::vector<int> snap_src, snap_dst;
  std::vector<int> lst, rst;

  for (int iter = 0; iter < max_iterations; ++iter) {
    bool any = false;
    for (int t = 0; t < T; ++t) any = any || active[t];
    if (!any) break;
    for (int t = 0; t < T; ++t) { attempt[t] = 0; accepted[t] = 0; searching[t] = 0; }
    // accepted steps of this iteration, applied to theta in one launch after
    // the line search (until then only the searching targets' rows are read)
    acc_dst.clear();
    acc_coef.clear();
    // Up to two attempts (CG direction, then steepest descent).
    for (int att = 0; att < 2; ++att) {
      lst.clear();
      for (int t = 0; t < T; ++t)
        if (active[t] && !accepted[t] && attempt[t] == att) lst.push_back(t);
      if (!dir_reduce_take(lst)) return fail_e("maxabs / dot");
      rst.clear();
      for (int t : lst) {
        dmax[t] = h_red[t];
        if (dmax[t] == 0.0) { searching[t] = 0; continue; }
        slope[t] = h_red[(size_t)T + t] / dmax[t];
        if (slope[t] >= 0.0 && !proxy) {
          // not a descent direction: restart along -grad (consumes this attempt)
          h_a[t] = -1.0;
          h_b[t] = 0.0;
          attempt[t] = att + 1;
          rst.push_back(t);
          continue;
        }
        searching[t] = 1;
        s_try[t] = step[t];
      }
      if (!dir_rows(grad, rst)) return fail_e("dir");
      // Batched Armijo backtracking (c1 = 1e-4, at most 25 trial steps per
      // target, halving after each failure).  One round evaluates the next
      // trial of every searching target; when few targets are left, each also
      // evaluates its following halvings in the same round (up to 4, never
      // more trial rows than targets) and takes the first step that passes --
      // the step sequential halving takes -- so stragglers cost fewer rounds.
      std::fill(used.begin(), used.end(), 0);
      for (;;) {
        which.clear();
        for (int t = 0; t < T; ++t)
          if (searching[t]) which.push_back(t);
        if (which.empty()) break;
        const int nb = std::max(1, std::min(4, T / (int)which.size()));
        prow.clear();
        pcoef.clear();
        pstep.clear();
        for (int t : which) {
          const int k = std::min(nb, 25 - used[t]);
          pfirst[t] = (int)prow.size();
          pcount[t] = k;
          double sj = s_try[t];
          for (int j = 0; j < k; ++j, sj *= 0.5) {
            prow.push_back(t);
            pcoef.push_back(sj / dmax[t]);
            pstep.push_back(sj);
          }
        }
        const int np = (int)prow.size();  // <= T
        const int* prow_d = (const int*)put(prow.data(), np * sizeof(int));
        const double* pcoef_d = (const double*)put(pcoef.data(), np * sizeof(double));
        if (!prow_d || !pcoef_d || !flush()) return fail_e("trial staging");
        // trial rows are read in place: theta + coef dir of their target
        ThetaRows trial;
        trial.theta = theta;
        trial.rows = prow_d;
        trial.dir = dir;
        trial.coef = pcoef_d;
        trial.stage = th_c;
        st = ilt_eval(p, np, targets, prow_d, trial, nullptr, nullptr, b, false);
        if (st != HUV_OK) return st;
        if (!get_costs(np, cost_c.data())) return fail_e("cost copy");
        for (int t : which) {
          bool ok = false;
          for (int j = pfirst[t]; j < pfirst[t] + pcount[t]; ++j) {
            const bool ok_j = proxy ? cost_c[j] < cost[t]
                                    : cost_c[j] <= cost[t] + 1e-4 * pstep[j] * slope[t];
            if (ok_j && std::isfinite(cost_c[j])) {
              searching[t] = 0;
              accepted[t] = 1;
              s_used[t] = pstep[j];
              acc_dst.push_back(t);
              acc_coef.push_back(pcoef[j]);
              ok = true;
              break;
            }
          }
          if (!ok) {
            used[t] += pcount[t];
            s_try[t] = pstep[pfirst[t] + pcount[t] - 1] * 0.5;
            if (used[t] >= 25) searching[t] = 0;  // exhausted
          }
        }
      }
      // Targets whose search failed: retry once along -grad with CG, else stop.
      rst.clear();
      for (int t = 0; t < T; ++t) {
        if (!active[t] || accepted[t]) continue;
        if (searching[t]) searching[t] = 0;  // exhausted backtracks
        if (attempt[t] == att && att == 0 && cg && dmax[t] != 0.0) {
          attempt[t] = 1;
          h_a[t] = -1.0;
          h_b[t] = 0.0;
          rst.push_back(t);
        }
      }
      if (att == 0 && !dir_rows(grad, rst)) return fail_e("dir");
    }
    bool any_acc = false;
    for (int t = 0; t < T; ++t) {
      if (!active[t]) continue;
      if (!accepted[t]) {
        active[t] = 0;
        term[t] = gnorm[t] > 0.0 ? 2 : 0;
      } else {
        any_acc = true;
      }
    }
    if (!any_acc) break;
    {
      // theta += coef dir in place for every accepted target, rounded exactly
      // as its trial point was
      const int* dr = (const int*)put(acc_dst.data(), acc_dst.size() * sizeof(int));
      const double* cf = (const double*)put(acc_coef.data(), acc_coef.size() * sizeof(double));
      if (!dr || !cf || !flush()) return fail_e("accept staging");
      e = huvk::launch_vec_axpy_rows(theta, dir, cf, N, (int)acc_dst.size(), dr, s);
      if (e != hipSuccess) return fail_e("accept copy");
    }
    which.clear();
    for (int t = 0; t < T; ++t)
      if (active[t] && accepted[t]) which.push_back(t);
    // the gradient evaluation's costs come with the CG reductions' synchronisation
    st = eval_sel(theta, gnew, cnew, which, true);
    if (st != HUV_OK) return st;
    // CG update: beta = max(0, (g'.g' - g'.g) / g.g)
    if (!reduce(grad, gnew, {AB, BB}, which)) return fail_e("dot");
    take_costs(cnew, which);
    const double* gg = gsq.data();
    const double* gng = h_red.data();
    const double* gngn = gng + T;
    for (int t : which) {
      const double prev = cost[t];
      cost[t] = cnew[t];
      it[t] += 1;
      cost_hist[(size_t)t * H + it[t]] = cost[t];
      gnorm[t] = std::sqrt(gngn[t]);
      if (tr.gnorm_hist) tr.gnorm_hist[(size_t)t * H + it[t]] = gnorm[t];
      const double beta = cg ? (gg[t] > 0.0 ? std::max(0.0, (gngn[t] - gng[t]) / gg[t]) : 0.0) : 0.0;
      gsq[t] = gngn[t];
      h_a[t] = -1.0;
      h_b[t] = beta;
      step[t] = std::min(s_used[t] * 1.5, 10.0 * learning_rate);
      const double rel = prev > 0.0 ? (prev - cost[t]) / prev : 0.0;
      small[t] = rel < tol ? small[t] + 1 : 0;
      if (small[t] >= 3 || cost[t] == 0.0) {
        active[t] = 0;
        term[t] = 0;
      }
      if (stop_cost && cost[t] <= stop_cost[t]) {
        active[t] = 0;
        term[t] = 3;
      }
      if (active[t] && it[t] >= limit[t]) active[t] = 0;  // MaxIterations
    }
    if (!dir_reduce_rows(gnew, which)) return fail_e("dir");
    // The new gradients become the current ones by exchanging the buffers:
    // every target still active next iteration is in `which`, so the rows
    // read later are all current; the other rows are never read again.
    std::swap(grad, gnew);
    // snapshots of the masks of targets whose global iteration is due
    snap_src.clear();
    snap_dst.clear();
    for (int t : which) {
      const int slot = slot_of(t);
      if (slot >= 0) {
        snap_src.push_back(t);
        snap_dst.push_back(t * tr.nslots + slot);
      }
    }
    if (!snap_src.empty()) {
      const int cnt = (int)snap_src.size();
      const int* sr = (const int*)put(snap_src.data(), cnt * sizeof(int));
      const int* dr = (const int*)put(snap_dst.data(), cnt * sizeof(int));
      if (!sr || !dr || !flush()) return fail_e("snapshot staging");
      ThetaRows snap;
      snap.theta = theta;
      snap.rows = sr;
      snap.stage = th_c;
      st = ilt_mask_t(p, cnt, snap, g_c, nullptr, b);
      if (st != HUV_OK) return st;
      e = huvk::launch_rows_scatter(g_c, dr, N, cnt, tr.snapshots, s);
      if (e != hipSuccess) return fail_e("snapshot scatter");
    }
  }
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail_e("synchronize");
  for (int t = 0; t < T; ++t) {
    iterations[t] = it[t];
    termination[t] = term[t];
  }
  return HUV_OK;
}

}  // namespace

extern "C" {

huv_status huv_ilt_optimize(huv_ilt* p, int T, const float* targets, float* theta,
                            int max_iterations, double learning_rate, double tol, int use_cg,
                            double* cost_hist, int* iterations, int* termination) {
  return ilt_optimize(p, T, targets, theta, max_iterations, learning_rate, tol, use_cg, nullptr,
                      cost_hist, iterations, termination);
}

huv_status huv_ilt_optimize_to(huv_ilt* p, int T, const float* targets, float* theta,
                               int max_iterations, double learning_rate, double tol, int use_cg,
                               const double* stop_cost, double* cost_hist, int* iterations,
                               int* termination) {
  if (!stop_cost) return fail(HUV_INVALID_PARAMETER, "null argument");
  return ilt_optimize(p, T, targets, theta, max_iterations, learning_rate, tol, use_cg, stop_cost,
                      cost_hist, iterations, termination);
}

huv_status huv_ilt_run_stage(huv_ilt* p, int T, const float* targets, float* theta,
                             int max_iterations, double learning_rate, double tol, int use_cg,
                             const double* stop_cost, double* cost_hist, double* gnorm_hist,
                             int snapshot_every, const int* iteration_offset, int nslots,
                             float* snapshots, int* iterations, int* termination) {
  StageTrace tr;
  tr.gnorm_hist = gnorm_hist;
  tr.every = snapshot_every;
  tr.offset = iteration_offset;
  tr.nslots = nslots;
  tr.snapshots = snapshots;
  return ilt_optimize(p, T, targets, theta, max_iterations, learning_rate, tol, use_cg, stop_cost,
                      cost_hist, iterations, termination, tr);
}

huv_status huv_ilt_set_binarization(huv_ilt* p, double bin_weight) {
  if (!p) return fail(HUV_INVALID_PARAMETER, "null argument");
  if (!(bin_weight >= 0.0 && std::isfinite(bin_weight)))
    return fail(HUV_INVALID_PARAMETER, "binarization_weight: must be non-negative and finite");
  std::lock_guard<std::mutex> lk(p->ctx->mu);
  p->cfg.bin_weight = bin_weight;
  return HUV_OK;
}

huv_status huv_ilt_mask(huv_ilt* p, int T, const float* theta, float* m) {
  if (!p || !theta || !m) return fail(HUV_INVALID_PARAMETER, "null argument");
  if (T < 1) return fail(HUV_INVALID_PARAMETER, "need at least one target");
  Ctx* c = p->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  hipSetDevice(c->device);
  IltBuffers b;
  huv_status st = ilt_buffers(c, p, T, b);
  if (st != HUV_OK) return st;
  ThetaRows src;
  src.theta = theta;
  return ilt_mask_t(p, T, src, m, nullptr, b);
}

}  // extern "C"
