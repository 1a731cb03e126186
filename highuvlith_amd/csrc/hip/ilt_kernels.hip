// HIP kernels of the ILT cost and exact adjoint gradient (gfx950).
//
// Reference semantics (crates/highuvlith-core/src/ilt.rs:28-45, 538-738):
//   m = sigma(beta theta), t = a + (1 - a) m           (a real)
//   I = flare(sum_k lambda_k |F^-1(K_k . F t)|^2), R = sigma(alpha (d I - t_r)) or d I
//   C = sum_c w_c (1/N) sum (R - Z)^2 + TV + binarisation
//   G = dC/dI, G0 = (1 - phi) G + (phi/N) sum G
//   dC/dm = 2 Re[(1 - a) sum_k lambda_k F^-1(K_k^* . F(G0 . A_k))] + reg;  dC/dtheta = dC/dm beta m (1 - m)
//
// Band-limited adjoint: F(G0 . A_k) is only needed on the kernel support
// |f| <= B, which only involves the |g| <= 2B band of G0; with G0 replaced by
// that band-limited G0_b the product G0_b A_k has per-axis degree 3B, so on
// a coarse grid m >= 4B + 1 its DFT is alias-free on |f| <= B.  The fused
// adjoint kernel therefore recomputes each A_k on the coarse grid, multiplies
// by G0_b sampled there, transforms back to the |f| <= B band and
// accumulates lambda_k conj(K_k) times it; only the band leaves the chip.
#include <hip/hip_runtime.h>

#include "fft_lds.hpp"
#include "ilt_kernels.hpp"

namespace huvk {

// m and t of evaluation rows read in place from the optimiser's vectors:
// row j's theta is theta[rows[j]] + coef[j] dir[rows[j]] (rows null: row j;
// coef null: theta alone), computed in FP64 and rounded as the line search's
// trial points are, so no compact copy of the rows is written.  m, t, tc may
// each be null; tc (complex absorber ar + i ai) replaces t when given.
// Grid: x over float4 pixel groups, y over rows.
__global__ void k_ilt_mask_rows(const float* __restrict__ theta, const int* __restrict__ rows,
                                const float* __restrict__ dir, const double* __restrict__ coef,
                                int64_t npix, float beta, float ar, float ai,
                                float* __restrict__ m, float* __restrict__ t,
                                float2* __restrict__ tc) {
  const int j = blockIdx.y;
  const size_t src = (size_t)(rows ? rows[j] : j) * npix, dst = (size_t)j * npix;
  const double c = coef ? coef[j] : 0.0;
  const int64_t n4 = npix / 4;
  const float4* th4 = reinterpret_cast<const float4*>(theta + src);
  const float4* d4 = coef ? reinterpret_cast<const float4*>(dir + src) : nullptr;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n4;
       i += (int64_t)gridDim.x * blockDim.x) {
    float v[4];
    const float4 a = th4[i];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    if (d4) {
      const float4 d = d4[i];
      v[0] = (float)((double)v[0] + c * (double)d.x);
      v[1] = (float)((double)v[1] + c * (double)d.y);
      v[2] = (float)((double)v[2] + c * (double)d.z);
      v[3] = (float)((double)v[3] + c * (double)d.w);
    }
    float mm[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) mm[k] = 1.0f / (1.0f + expf(-beta * v[k]));
    if (m) reinterpret_cast<float4*>(m + dst)[i] = make_float4(mm[0], mm[1], mm[2], mm[3]);
    if (tc) {
      float4* o = reinterpret_cast<float4*>(tc + dst);
      o[2 * i] = make_float4(ar + (1.0f - ar) * mm[0], ai - ai * mm[0], ar + (1.0f - ar) * mm[1],
                             ai - ai * mm[1]);
      o[2 * i + 1] = make_float4(ar + (1.0f - ar) * mm[2], ai - ai * mm[2],
                                 ar + (1.0f - ar) * mm[3], ai - ai * mm[3]);
    } else if (t) {
      reinterpret_cast<float4*>(t + dst)[i] =
          make_float4(ar + (1.0f - ar) * mm[0], ar + (1.0f - ar) * mm[1], ar + (1.0f - ar) * mm[2],
                      ar + (1.0f - ar) * mm[3]);
    }
  }
}

// Gaussian density filter on the band (ilt.rs:443-454): the Hermitian half
// band Ib[t][gy + Bf][gx] (gx in [0, Bf]) = X[gx][gy] exp(-c (gx^2 + gy^2)) / n^2
// of the band box X[t][gx + Bf][gy + Bf] of a real field (band_image input).
__global__ void k_ilt_filter_band(const float2* __restrict__ X, int Bf, float c, float inv_n2,
                                  int count, float2* __restrict__ Ib) {
  const int W = 2 * Bf + 1, half = (Bf + 1) * W;
  const int64_t total = (int64_t)count * half;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = i / half;
    const int r = (int)(i - t * half);
    const int gy = r / (Bf + 1) - Bf, gx = r % (Bf + 1);
    const float h = expf(-c * (float)(gx * gx + gy * gy)) * inv_n2;
    const float2 v = X[(size_t)t * W * W + (size_t)(gx + Bf) * W + (gy + Bf)];
    Ib[i] = make_float2(v.x * h, v.y * h);
  }
}

// Block-wise partial sums in FP64, fixed order within the block.
template <int THREADS>
__device__ __forceinline__ double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int s = THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

constexpr int PW_THREADS = 256;

// (The fidelity term of each image and G = dC/dI are evaluated by the
// band-image stage itself, launch_band_image_fidelity.)

// Regularisers on m (ilt.rs:606-645): cost partials per block and gradient.
__global__ __launch_bounds__(PW_THREADS) void k_ilt_reg(
    const float* __restrict__ m, int n, int blocks_per_img, float tv_w, float tv_eps, float bin_w,
    float* __restrict__ reg_grad, double* __restrict__ partial) {
  __shared__ double sh[PW_THREADS];
  const int tgt = blockIdx.y;
  const int64_t npix = (int64_t)n * n;
  const float* mm = m + (size_t)tgt * npix;
  float* rg = reg_grad + (size_t)tgt * npix;
  const float inv_n = 1.0f / (float)npix;
  double acc = 0.0;
  const int64_t chunk = (npix + blocks_per_img - 1) / blocks_per_img;
  const int64_t lo = blockIdx.x * chunk, hi = min(npix, lo + chunk);
  auto sval = [&](int i, int j, float& gx, float& gy) {
    const float v = mm[(size_t)i * n + j];
    gx = mm[(size_t)i * n + (j + 1) % n] - v;
    gy = mm[(size_t)((i + 1) % n) * n + j] - v;
    return sqrtf(gx * gx + gy * gy + tv_eps * tv_eps);
  };
  for (int64_t p = lo + threadIdx.x; p < hi; p += PW_THREADS) {
    const int i = (int)(p / n), j = (int)(p % n);
    float g = 0.0f;
    if (tv_w > 0.0f) {
      float gx, gy, lx, ly, ux, uy;
      const float s0 = sval(i, j, gx, gy);
      const float sl = sval(i, (j + n - 1) % n, lx, ly);
      const float su = sval((i + n - 1) % n, j, ux, uy);
      acc += (double)(tv_w * (s0 - tv_eps) * inv_n);
      g += tv_w * (-(gx + gy) / s0 + lx / sl + uy / su) * inv_n;
    }
    if (bin_w > 0.0f) {
      const float v = mm[p];
      acc += (double)(bin_w * 4.0f * v * (1.0f - v) * inv_n);
      g += bin_w * 4.0f * (1.0f - 2.0f * v) * inv_n;
    }
    rg[p] = g;
  }
  const double s = block_sum<PW_THREADS>(acc, sh);
  if (threadIdx.x == 0) partial[(size_t)tgt * blocks_per_img + blockIdx.x] = s;
}

// cost[t] = sum_c w_c (1/N) sum_blocks fid[t, c] + sum_blocks reg[t]  (fixed order).
__global__ void k_ilt_cost_reduce(const double* __restrict__ fid, const double* __restrict__ reg,
                                  int ntargets, int ncond, int bpi, int reg_bpi,
                                  const float* __restrict__ weight, double inv_n,
                                  double* __restrict__ cost) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntargets) return;
  double total = 0.0;
  for (int c = 0; c < ncond; ++c) {
    double s = 0.0;
    const double* p = fid + ((size_t)t * ncond + c) * bpi;
    for (int b = 0; b < bpi; ++b) s += p[b];
    total += (double)weight[c] * s * inv_n;
  }
  if (reg) {
    const double* p = reg + (size_t)t * reg_bpi;
    for (int b = 0; b < reg_bpi; ++b) total += p[b];
  }
  cost[t] = total;
}

// ---------------------------------------------------------------------------
// Adjoint, coarse grid M <= 64 (a single x class), in two launches:
//   k_ilt_g0:      one work-group per image (target, condition): G0_b on the
//                  coarse grid, G0[img][y][x] (real), from the |g| <= 2B band
//                  Gb of F(G), layout [gx + 2B][gy + 2B], flare adjoint applied.
//   k_ilt_adjoint: work-group (image, split s) takes the kernels k = s, s + KS,
//                  ...: field A_k, product with G0_b, back to the |f| <= B band,
//                  accumulating lambda_k conj(K_k) F(G0_b . A_k) / n^2 into its
//                  own partial band Rpart[img][s][fx + B][fy + B]; k_ilt_hband
//                  sums the KS partials in a fixed order.
// ---------------------------------------------------------------------------
template <int M>
__global__ __launch_bounds__(M) void k_ilt_g0(const float2* __restrict__ Gb, int B, float g0_scale,
                                              float flare, float* __restrict__ G0g) {
  constexpr int TP = M + 1;
  extern __shared__ unsigned char smem_raw[];
  c32* T = reinterpret_cast<c32*>(smem_raw);  // (4B+1) x TP
  const int G = 2 * B, W2 = 4 * B + 1;
  const int img = blockIdx.x;
  const int tid = threadIdx.x;
  const float2* gb = Gb + (size_t)img * W2 * W2;
  // x pass over the 4B+1 band rows gy
  for (int iy = tid; iy < W2; iy += M) {
    c32 v[M];
#pragma unroll
    for (int h = 0; h < M; ++h) {
      v[h] = c32{0.f, 0.f};
      // gx = h (h <= 2B) or h - M (>= -2B)
      if (h <= G) {
        const float2 a = gb[(size_t)(h + G) * W2 + iy];
        v[h] = c32{a.x, a.y};
      }
      if (h - M >= -G) {
        const float2 a = gb[(size_t)(h - M + G) * W2 + iy];
        v[h] = cadd(v[h], c32{a.x, a.y});
      }
    }
    // flare adjoint: non-DC terms scale by (1 - phi)
    const int gy = iy - G;
#pragma unroll
    for (int h = 0; h < M; ++h)
      if (!(h == 0 && gy == 0)) v[h] = cscale(v[h], 1.0f - flare);
    fft_reg<+1, M>(v);
#pragma unroll
    for (int j = 0; j < M; ++j) T[iy * TP + j] = v[j];
  }
  __syncthreads();
  const int j = tid;
  c32 d[M];
#pragma unroll
  for (int h = 0; h < M; ++h) {
    d[h] = c32{0.f, 0.f};
    if (h <= G) d[h] = T[(h + G) * TP + j];
    if (h - M >= -G) d[h] = cadd(d[h], T[(h - M + G) * TP + j]);
  }
  fft_reg<+1, M>(d);
  float* g0 = G0g + (size_t)img * M * M;
#pragma unroll
  for (int y = 0; y < M; ++y) g0[y * M + j] = d[y].re * g0_scale;
}

template <int M>
__global__ __launch_bounds__(M) void k_ilt_adjoint(
    const float2* __restrict__ Mq, const float* __restrict__ G0g, int B, int ncond,
    const int* __restrict__ cond_K, const float2* const* __restrict__ cond_kern,
    const float* const* __restrict__ cond_lam, float lam_scale, float2* __restrict__ Rpart) {
  constexpr int TP = M + 1;
  extern __shared__ unsigned char smem_raw[];
  const int W = 2 * B + 1;
  c32* T = reinterpret_cast<c32*>(smem_raw);                  // W x TP
  c32* P1 = T + (size_t)W * TP;                               // W x TP
  float* G0 = reinterpret_cast<float*>(P1 + (size_t)W * TP);  // M x TP
  const int img = blockIdx.x, split = blockIdx.y, KS = gridDim.y;
  const int tgt = img / ncond, c = img % ncond;
  const int tid = threadIdx.x;
  const float2* mq = Mq + (size_t)tgt * W * W;
  {
    const float* g0 = G0g + (size_t)img * M * M;
    for (int y = 0; y < M; ++y) G0[y * TP + tid] = g0[y * M + tid];
  }
  __syncthreads();

  // ---- per kernel: field, product, back to the |f| <= B band ----
  const int K = cond_K[c];
  const float2* kern = cond_kern[c];
  const float* lam = cond_lam[c];
  c32 R[2 * 15 + 1];  // accumulator row fy = tid - B (B <= 15)
#pragma unroll
  for (int i = 0; i < 31; ++i) R[i] = c32{0.f, 0.f};
  for (int k = split; k < K; k += KS) {
    // x pass (rows iy in [0, 2B]): E row = IFFT_M(K . M)
    for (int iy = tid; iy < W; iy += M) {
      const float2* kk = kern + (size_t)k * W * W + iy;
      const float2* mm = mq + iy;
      c32 v[M];
#pragma unroll
      for (int h = 0; h < M; ++h) {
        v[h] = c32{0.f, 0.f};
        if (h <= B) {
          const float2 a = kk[(size_t)(h + B) * W], b = mm[(size_t)(h + B) * W];
          v[h] = cmul(c32{a.x, a.y}, c32{b.x, b.y});
        }
        if (h - M >= -B) {
          const float2 a = kk[(size_t)(h - M + B) * W], b = mm[(size_t)(h - M + B) * W];
          v[h] = cmac(v[h], c32{a.x, a.y}, c32{b.x, b.y});
        }
      }
      fft_reg<+1, M>(v);
#pragma unroll
      for (int j = 0; j < M; ++j) T[iy * TP + j] = v[j];
    }
    __syncthreads();
    // y pass (column j): E column, times G0_b, forward FFT_M, keep |fy| <= B
    {
      const int j = tid;
      c32 d[M];
#pragma unroll
      for (int h = 0; h < M; ++h) {
        d[h] = c32{0.f, 0.f};
        if (h <= B) d[h] = T[(h + B) * TP + j];
        if (h - M >= -B) d[h] = cadd(d[h], T[(h - M + B) * TP + j]);
      }
      fft_reg<+1, M>(d);
#pragma unroll
      for (int y = 0; y < M; ++y) d[y] = cscale(d[y], G0[y * TP + j]);
      fft_reg<-1, M>(d);
#pragma unroll
      for (int h = 0; h < M; ++h) {
        if (h <= B) P1[(h + B) * TP + j] = d[h];
        if (h - M >= -B) P1[(h - M + B) * TP + j] = d[h];
      }
    }
    __syncthreads();
    // x stage (row fy): forward FFT_M over x, keep |fx| <= B, accumulate
    if (tid < W) {
      c32 v[M];
#pragma unroll
      for (int j = 0; j < M; ++j) v[j] = P1[tid * TP + j];
      fft_reg<-1, M>(v);
      const float l = lam[k] * lam_scale;
      const float2* kk = kern + (size_t)k * W * W + tid;  // K[k][fx + B][fy + B], fy = tid - B
#pragma unroll
      for (int i = 0; i < 31; ++i) {
        const int fx = i - 15;
        if (fx >= -B && fx <= B) {
          const c32 val = v[(fx + M) % M];
          const float2 kv = kk[(size_t)(fx + B) * W];
          R[i] = cadd(R[i], cscale(cconjmul(c32{kv.x, kv.y}, val), l));
        }
      }
    }
    __syncthreads();
  }
  if (tid < W) {
    float2* ro = Rpart + ((size_t)img * KS + split) * W * W + tid;
#pragma unroll
    for (int i = 0; i < 31; ++i) {
      const int fx = i - 15;
      if (fx >= -B && fx <= B) ro[(size_t)(fx + B) * W] = make_float2(R[i].re, R[i].im);
    }
  }
}

// ---------------------------------------------------------------------------
// Adjoint on coarse grids M = 128, 192, 256 (band half-width B up to 63): the
// same algebra as k_ilt_g0 / k_ilt_adjoint, with the length-M transforms done
// by teams of threads (four-step, LDS) and the per-(image, kernel) field held
// in HBM between passes.  A job is (image, kernel) with image = target x
// ncond + condition and kernel k < KS (KS = the largest kernel count; jobs
// with k >= K_c of their condition contribute zero).
//   g0_rows / g0_cols: G0_b on the coarse grid (as k_ilt_g0), via tmp
//                      [img][4B+1][M].
//   adj_rows:  band row ky: T[job][ky][x] = IFFT_M(K_k . M)(x).
//   adj_cols:  column x: E = IFFT_M over ky, times G0_b[y][x], FFT_M back,
//              rows |fy| <= B stored in place of T.
//   adj_out:   row fy: FFT_M over x, |fx| <= B kept, times lambda_k conj(K_k)
//              into Rpart[img][k][fx + B][fy + B] (summed by k_ilt_hband).
// ---------------------------------------------------------------------------
template <int M>
struct AdjFft {
  static constexpr int N1 = M == 128 ? 8 : M == 192 ? 12 : 16;
  static constexpr int N2 = M / N1;
  static constexpr int TEAMS = 256 / N2;
};
constexpr int kAdjColTeams = 8;

// signed frequency of FFT bin h if |f| <= band, else a value outside the band
__device__ __forceinline__ int band_freq(int h, int M, int band) {
  if (h <= band) return h;
  if (h - M >= -band) return h - M;
  return 1 << 20;
}

template <int M>
__global__ __launch_bounds__(256) void k_ilt_g0_rows(const float2* __restrict__ Gb, int B,
                                                     float flare, const float2* __restrict__ tw_m,
                                                     float2* __restrict__ tmp) {
  using F = AdjFft<M>;
  __shared__ c32 scratch[F::TEAMS][(F::N1 + 1) * F::N2];
  const int team = threadIdx.x / F::N2, t = threadIdx.x % F::N2;
  const int G = 2 * B, W2 = 4 * B + 1;
  const int img = blockIdx.y;
  const int iy = blockIdx.x * F::TEAMS + team;
  const bool active = iy < W2;
  const float2* gb = Gb + (size_t)img * W2 * W2;
  const int gy = iy - G;
  fft_four_step<+1, F::N1, F::N2>(
      t, scratch[team], tw_m,
      [&](int h) {
        const int gx = band_freq(h, M, G);
        if (!active || gx > G) return c32{0.f, 0.f};
        const float2 a = gb[(size_t)(gx + G) * W2 + iy];
        // flare adjoint: non-DC terms scale by (1 - phi)
        const float f = (gx == 0 && gy == 0) ? 1.0f : 1.0f - flare;
        return c32{a.x * f, a.y * f};
      },
      [&](int x, c32 v) {
        if (active) tmp[((size_t)img * W2 + iy) * M + x] = make_float2(v.re, v.im);
      });
}

template <int M>
__global__ __launch_bounds__(256) void k_ilt_g0_cols(const float2* __restrict__ tmp, int B,
                                                     float g0_scale, const float2* __restrict__ tw_m,
                                                     float* __restrict__ G0g) {
  using F = AdjFft<M>;
  __shared__ c32 scratch[F::TEAMS][(F::N1 + 1) * F::N2];
  const int team = threadIdx.x / F::N2, t = threadIdx.x % F::N2;
  const int G = 2 * B, W2 = 4 * B + 1;
  const int img = blockIdx.y;
  const int x = blockIdx.x * F::TEAMS + team;
  const bool active = x < M;
  fft_four_step<+1, F::N1, F::N2>(
      t, scratch[team], tw_m,
      [&](int h) {
        const int gy = band_freq(h, M, G);
        if (!active || gy > G) return c32{0.f, 0.f};
        const float2 a = tmp[((size_t)img * W2 + gy + G) * M + x];
        return c32{a.x, a.y};
      },
      [&](int y, c32 v) {
        if (active) G0g[((size_t)img * M + y) * M + x] = v.re * g0_scale;
      });
}

template <int M>
__global__ __launch_bounds__(256) void k_ilt_adj_rows(
    const float2* __restrict__ Mq, int B, int ncond, int KS, const int* __restrict__ cond_K,
    const float2* const* __restrict__ cond_kern, const float2* __restrict__ tw_m,
    float2* __restrict__ buf) {
  using F = AdjFft<M>;
  __shared__ c32 scratch[F::TEAMS][(F::N1 + 1) * F::N2];
  const int team = threadIdx.x / F::N2, t = threadIdx.x % F::N2;
  const int W = 2 * B + 1;
  const int job = blockIdx.y, img = job / KS, k = job % KS;
  const int tgt = img / ncond, c = img % ncond;
  const int iy = blockIdx.x * F::TEAMS + team;
  const bool active = iy < W;
  const bool live = active && k < cond_K[c];
  const float2* kk = cond_kern[c] + (size_t)(live ? k : 0) * W * W;
  const float2* mq = Mq + (size_t)tgt * W * W;
  fft_four_step<+1, F::N1, F::N2>(
      t, scratch[team], tw_m,
      [&](int h) {
        const int kx = band_freq(h, M, B);
        if (!live || kx > B) return c32{0.f, 0.f};
        const float2 a = kk[(size_t)(kx + B) * W + iy], b = mq[(size_t)(kx + B) * W + iy];
        return cmul(c32{a.x, a.y}, c32{b.x, b.y});
      },
      [&](int x, c32 v) {
        if (active) buf[((size_t)job * W + iy) * M + x] = make_float2(v.re, v.im);
      });
}

template <int M>
__global__ __launch_bounds__(kAdjColTeams * AdjFft<M>::N2) void k_ilt_adj_cols(
    const float* __restrict__ G0g, int B, int KS, const float2* __restrict__ tw_m,
    float2* __restrict__ buf) {
  using F = AdjFft<M>;
  // fewer teams than the other passes: each also holds its column (M values)
  __shared__ c32 scratch[kAdjColTeams][(F::N1 + 1) * F::N2];
  __shared__ c32 col[kAdjColTeams][M];
  const int team = threadIdx.x / F::N2, t = threadIdx.x % F::N2;
  const int W = 2 * B + 1;
  const int job = blockIdx.y, img = job / KS;
  const int x = blockIdx.x * kAdjColTeams + team;
  const bool active = x < M;
  const float* g0 = G0g + (size_t)img * M * M;
  float2* b = buf + (size_t)job * W * M;
  // E[y][x] = IFFT over ky, times G0_b[y][x]
  fft_four_step<+1, F::N1, F::N2>(
      t, scratch[team], tw_m,
      [&](int h) {
        const int ky = band_freq(h, M, B);
        if (!active || ky > B) return c32{0.f, 0.f};
        const float2 a = b[(size_t)(ky + B) * M + x];
        return c32{a.x, a.y};
      },
      [&](int y, c32 v) { col[team][y] = active ? cscale(v, g0[(size_t)y * M + x]) : c32{0.f, 0.f}; });
  // forward FFT over y; rows |fy| <= B replace this column of T
  fft_four_step<-1, F::N1, F::N2>(
      t, scratch[team], tw_m, [&](int y) { return col[team][y]; },
      [&](int f, c32 v) {
        const int fy = band_freq(f, M, B);
        if (active && fy <= B) b[(size_t)(fy + B) * M + x] = make_float2(v.re, v.im);
      });
}

template <int M>
__global__ __launch_bounds__(256) void k_ilt_adj_out(
    const float2* __restrict__ buf, int B, int ncond, int KS, const int* __restrict__ cond_K,
    const float2* const* __restrict__ cond_kern, const float* const* __restrict__ cond_lam,
    float lam_scale, const float2* __restrict__ tw_m, float2* __restrict__ Rpart) {
  using F = AdjFft<M>;
  __shared__ c32 scratch[F::TEAMS][(F::N1 + 1) * F::N2];
  const int team = threadIdx.x / F::N2, t = threadIdx.x % F::N2;
  const int W = 2 * B + 1;
  const int job = blockIdx.y, img = job / KS, k = job % KS;
  const int c = img % ncond;
  const int fyi = blockIdx.x * F::TEAMS + team;  // fy + B
  const bool active = fyi < W;
  const bool live = active && k < cond_K[c];
  const float2* kk = cond_kern[c] + (size_t)(live ? k : 0) * W * W;
  const float l = live ? cond_lam[c][k] * lam_scale : 0.0f;
  const float2* row = buf + ((size_t)job * W + (active ? fyi : 0)) * M;
  float2* ro = Rpart + (size_t)job * W * W;
  fft_four_step<-1, F::N1, F::N2>(
      t, scratch[team], tw_m,
      [&](int x) {
        if (!live) return c32{0.f, 0.f};
        const float2 a = row[x];
        return c32{a.x, a.y};
      },
      [&](int f, c32 v) {
        const int fx = band_freq(f, M, B);
        if (!active || fx > B) return;
        c32 r{0.f, 0.f};
        if (live) {
          const float2 kv = kk[(size_t)(fx + B) * W + fyi];
          r = cscale(cconjmul(c32{kv.x, kv.y}, v), l);
        }
        ro[(size_t)(fx + B) * W + fyi] = make_float2(r.re, r.im);
      });
}

// H[fy][fx] (fx in [0, B], fy in [-B, B]) = sum_c w_c (1-a) (R_c[f] + conj R_c[-f]) / n^2,
// layout [fy + B][fx] (the Hermitian half band read by band_image, half-width B).
// R_c = sum over the KS partial bands of image (tgt, c), in split order.
// With a complex absorber a and c = 1 - a (ilt.rs:672,715), 2 Re[conj(c) Q]
// has the spectrum conj(c) R(f) + c conj(R(-f)); inv_n2 = 1 / n^2.
__global__ void k_ilt_hband(const float2* __restrict__ Rin, int B, int ncond, int KS,
                            const float* __restrict__ weight, float cr, float ci, float inv_n2,
                            float2* __restrict__ H) {
  const int tgt = blockIdx.y;
  const int W = 2 * B + 1;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (B + 1) * W) return;
  const int fx = idx % (B + 1), fy = idx / (B + 1) - B;
  c32 acc{0.f, 0.f};
  for (int c = 0; c < ncond; ++c) {
    c32 p{0.f, 0.f}, q{0.f, 0.f};
    for (int sp = 0; sp < KS; ++sp) {
      const float2* r = Rin + (((size_t)tgt * ncond + c) * KS + sp) * W * W;
      const float2 a = r[(size_t)(fx + B) * W + (fy + B)];
      const float2 b = r[(size_t)(B - fx) * W + (B - fy)];
      p = cadd(p, c32{a.x, a.y});
      q = cadd(q, c32{b.x, b.y});
    }
    // conj(c) p + c conj(q)
    const c32 u = cconjmul(c32{cr, ci}, p);
    const c32 v = cmul(c32{cr, ci}, c32{q.re, -q.im});
    acc = cadd(acc, cscale(cadd(u, v), weight[c]));
  }
  H[(size_t)tgt * (B + 1) * W + idx] = make_float2(acc.re * inv_n2, acc.im * inv_n2);
}

// Local proxy gradient (ilt.rs:719-721): dcdm[t] = sum_c w_c G0_c with
// G0 = (1 - phi) G + (phi / N) sum G; gsum[img] are the per-image sums of G
// (k_ilt_gsum partials, summed here in block order).
__global__ __launch_bounds__(256) void k_ilt_gsum(const float* __restrict__ G, int64_t npix,
                                                  int bpi, double* __restrict__ part) {
  __shared__ double sh[256];
  const int img = blockIdx.y;
  const float* g = G + (size_t)img * npix;
  const int64_t chunk = (npix + bpi - 1) / bpi;
  const int64_t lo = blockIdx.x * chunk, hi = min(npix, lo + chunk);
  double acc = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += 256) acc += (double)g[i];
  const double s = block_sum<256>(acc, sh);
  if (threadIdx.x == 0) part[(size_t)img * bpi + blockIdx.x] = s;
}

__global__ void k_ilt_proxy(const float* __restrict__ G, const double* __restrict__ part,
                            int ntargets, int ncond, int64_t npix, int bpi,
                            const float* __restrict__ weight, float flare,
                            float* __restrict__ dcdm) {
  const int64_t total = (int64_t)ntargets * npix;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = i / npix, p = i - t * npix;
    float acc = 0.0f;
    for (int c = 0; c < ncond; ++c) {
      const size_t img = (size_t)t * ncond + c;
      double s = 0.0;
      for (int b = 0; b < bpi; ++b) s += part[img * bpi + b];
      const float g0 = (1.0f - flare) * G[img * npix + p] + flare * (float)(s / (double)npix);
      acc += weight[c] * g0;
    }
    dcdm[i] = acc;
  }
}

// grad = (dcdm + reg) * beta m (1 - m).
// grad row orows[j] (null: j) of evaluation row j.  Grid: x over pixels, y over rows.
__global__ void k_ilt_chain(const float* __restrict__ dcdm, const float* __restrict__ reg,
                            const float* __restrict__ m, float beta, int64_t npix,
                            const int* __restrict__ orows, float* __restrict__ grad) {
  const int j = blockIdx.y;
  const size_t a = (size_t)j * npix, o = (size_t)(orows ? orows[j] : j) * npix;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < npix;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float mm = m[a + i];
    const float r = reg ? reg[a + i] : 0.0f;
    grad[o + i] = (dcdm[a + i] + r) * beta * mm * (1.0f - mm);
  }
}

// ---------------------------------------------------------------------------
// Batched vector operations of the CG / Armijo loop (ilt.rs:755-866).
// Per-target FP64 partial sums, reduced in a fixed order.
// ---------------------------------------------------------------------------
// out[t * nblk + b] = partial of sum_i x_i y_i  (mode 0) or max |x_i| (mode 1)
// Per-target vector operations over a list of target rows (rows = null: all
// rows 0..count-1).  Grid: x over pixel blocks, y over the listed rows.
__device__ __forceinline__ int row_of(const int* __restrict__ rows, int j) {
  return rows ? rows[j] : j;
}

// Up to three per-target reductions over two vectors a, b in one pass; op r
// (4 bits of ops each, r < nops): 0 a.a, 1 a.b, 2 b.b, 3 max |a|.  Each
// reduction keeps the fixed order of a single one: per-thread FP64 sums over
// the block's strided pixels, then a tree over the block.
// part[(r * T + t) * nblk + block].
// acc after one more pixel of reduction op, each product and sum rounded on
// its own (no fused multiply-add), whatever the caller's contraction setting.
__device__ __forceinline__ double reduce_step(int op, double acc, float av, float bv) {
#pragma clang fp contract(off)
  switch (op) {
    case 0: return acc + (double)av * (double)av;
    case 1: return acc + (double)av * (double)bv;
    case 2: return acc + (double)bv * (double)bv;
    default: return fmax(acc, (double)fabsf(av));
  }
}

// Pixels per thread whose loads are issued together before their terms are
// added (in pixel order, so the sums are those of one pixel at a time).
constexpr int kReduceBatch = 8;

template <int NOPS>
__global__ __launch_bounds__(PW_THREADS) void k_vec_reduce(const float* __restrict__ a,
                                                           const float* __restrict__ b,
                                                           int64_t npix, int nblk, int ops, int T,
                                                           const int* __restrict__ rows,
                                                           double* __restrict__ part) {
  __shared__ double sh[NOPS][PW_THREADS];
  const int t = row_of(rows, blockIdx.y);
  const float* aa = a + (size_t)t * npix;
  const float* bb = b ? b + (size_t)t * npix : nullptr;
  const int64_t chunk = (npix + nblk - 1) / nblk;
  const int64_t lo = blockIdx.x * chunk, hi = min(npix, lo + chunk);
  int op[NOPS];
#pragma unroll
  for (int r = 0; r < NOPS; ++r) op[r] = (ops >> (4 * r)) & 15;
  double acc[NOPS];
#pragma unroll
  for (int r = 0; r < NOPS; ++r) acc[r] = 0.0;
  constexpr int U = kReduceBatch;
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += (int64_t)U * PW_THREADS) {
    float av[U], bv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      // past the block's end: a valid pixel, loaded and then ignored
      const int64_t i = min(i0 + (int64_t)u * PW_THREADS, hi - 1);
      av[u] = aa[i];
      bv[u] = bb ? bb[i] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (i0 + (int64_t)u * PW_THREADS >= hi) break;
#pragma unroll
      for (int r = 0; r < NOPS; ++r) acc[r] = reduce_step(op[r], acc[r], av[u], bv[u]);
    }
  }
#pragma unroll
  for (int r = 0; r < NOPS; ++r) sh[r][threadIdx.x] = acc[r];
  __syncthreads();
  for (int s = PW_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int r = 0; r < NOPS; ++r)
        sh[r][threadIdx.x] = op[r] == 3 ? fmax(sh[r][threadIdx.x], sh[r][threadIdx.x + s])
                                        : sh[r][threadIdx.x] + sh[r][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x < NOPS)
    part[((size_t)threadIdx.x * T + t) * nblk + blockIdx.x] = sh[threadIdx.x][0];
}

// out[r * T + t] = sum (or max) over the blocks of part, for the listed rows.
__global__ void k_vec_finish(const double* __restrict__ part, int count, int nblk, int ops,
                             int nops, int T, const int* __restrict__ rows,
                             double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (j >= count || r >= nops) return;
  const int t = row_of(rows, j);
  const bool mx = ((ops >> (4 * r)) & 15) == 3;
  const double* p = part + ((size_t)r * T + t) * nblk;
  double acc = 0.0;
  for (int k = 0; k < nblk; ++k) acc = mx ? fmax(acc, p[k]) : acc + p[k];
  out[(size_t)r * T + t] = acc;
}

// d = a[t] * g + b[t] * d
__global__ void k_vec_dir(const float* __restrict__ g, const double* __restrict__ a,
                          const double* __restrict__ b, int64_t npix, const int* __restrict__ rows,
                          float* __restrict__ d) {
  const int t = row_of(rows, blockIdx.y);
  const size_t base = (size_t)t * npix;
  const double ca = a[t], cb = b[t];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < npix;
       i += (int64_t)gridDim.x * blockDim.x)
    d[base + i] = (float)(ca * (double)g[base + i] + cb * (double)d[base + i]);
}

// k_vec_dir fused with the two reductions that follow a direction update,
// max |d| and d.g of the new d: the grid, pixel order and block tree of
// k_vec_reduce<2> with ops (3, 1), so part[] holds exactly what that
// reduction would write (part[(r * T + t) * nblk + block], r = 0 max |d|,
// r = 1 d.g).
__global__ __launch_bounds__(PW_THREADS) void k_vec_dir_reduce(
    const float* __restrict__ g, const double* __restrict__ a, const double* __restrict__ b,
    int64_t npix, int nblk, int T, const int* __restrict__ rows, float* __restrict__ d,
    double* __restrict__ part) {
  __shared__ double sh[2][PW_THREADS];
  const int t = row_of(rows, blockIdx.y);
  const size_t base = (size_t)t * npix;
  const double ca = a[t], cb = b[t];
  const int64_t chunk = (npix + nblk - 1) / nblk;
  const int64_t lo = blockIdx.x * chunk, hi = min(npix, lo + chunk);
  double acc0 = 0.0, acc1 = 0.0;
  constexpr int U = kReduceBatch;
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += (int64_t)U * PW_THREADS) {
    // The rounding of each step is spelled out: the update as k_vec_dir
    // compiles it (cb d rounded, then one fused multiply-add with ca g), the
    // reductions as reduce_step.
#pragma clang fp contract(off)
    float gv[U], dold[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = min(i0 + (int64_t)u * PW_THREADS, hi - 1);
      gv[u] = g[base + i];
      dold[u] = d[base + i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = i0 + (int64_t)u * PW_THREADS;
      if (i >= hi) break;
      const double cbd = cb * (double)dold[u];
      const float dv = (float)fma(ca, (double)gv[u], cbd);
      d[base + i] = dv;
      acc0 = reduce_step(3, acc0, dv, gv[u]);
      acc1 = reduce_step(1, acc1, dv, gv[u]);
    }
  }
  sh[0][threadIdx.x] = acc0;
  sh[1][threadIdx.x] = acc1;
  __syncthreads();
  for (int s = PW_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sh[0][threadIdx.x] = fmax(sh[0][threadIdx.x], sh[0][threadIdx.x + s]);
      sh[1][threadIdx.x] = sh[1][threadIdx.x] + sh[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x < 2)
    part[((size_t)threadIdx.x * T + t) * nblk + blockIdx.x] = sh[threadIdx.x][0];
}

// Line-search trials: out[j] = x[rows[j]] + coef[j] d[rows[j]] (compact rows).
__global__ void k_vec_axpy_gather(const float* __restrict__ x, const float* __restrict__ d,
                                  const double* __restrict__ coef, int64_t npix,
                                  const int* __restrict__ rows, float* __restrict__ out) {
  const int j = blockIdx.y;
  const size_t src = (size_t)rows[j] * npix, dst = (size_t)j * npix;
  const double c = coef[j];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < npix;
       i += (int64_t)gridDim.x * blockDim.x)
    out[dst + i] = (float)((double)x[src + i] + c * (double)d[src + i]);
}

// Accepted line-search steps, in place: x[rows[j]] += coef[j] d[rows[j]], in
// FP64 and rounded exactly as the trial point was (k_ilt_mask_rows).
__global__ void k_vec_axpy_rows(float* __restrict__ x, const float* __restrict__ d,
                                const double* __restrict__ coef, int64_t npix,
                                const int* __restrict__ rows) {
  const int j = blockIdx.y;
  const size_t r = (size_t)rows[j] * npix;
  const double c = coef[j];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < npix;
       i += (int64_t)gridDim.x * blockDim.x)
    x[r + i] = (float)((double)x[r + i] + c * (double)d[r + i]);
}

// Row gather / scatter of per-target images: dst[j] = src[idx[j]] or
// dst[idx[j]] = src[j], j < nrows (rows of npix floats).
template <bool SCATTER>
__global__ void k_rows_move(const float* __restrict__ src, const int* __restrict__ idx,
                            int64_t npix, int nrows, float* __restrict__ dst) {
  const int64_t total = npix * nrows;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = i / npix, k = i - j * npix;
    const int64_t r = idx[j];
    if (SCATTER) dst[r * npix + k] = src[i];
    else dst[i] = src[r * npix + k];
  }
}

// ===========================================================================
// Launchers
// ===========================================================================
static int grid_for(int64_t total, int threads) {
  const int64_t b = (total + threads - 1) / threads;
  return (int)std::min<int64_t>(b, 1 << 16);
}

hipError_t launch_ilt_mask_rows(const float* theta, const int* rows, const float* dir,
                                const double* coef, int64_t npix, int count, float beta, float ar,
                                float ai, float* m, float* t, float2* tc, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (npix % 4 != 0 || (coef && !dir)) return hipErrorInvalidValue;
  const int64_t bx = std::min<int64_t>((npix / 4 + 255) / 256, 1024);
  hipLaunchKernelGGL(k_ilt_mask_rows, dim3((unsigned)bx, (unsigned)count), dim3(256), 0, s, theta,
                     rows, dir, coef, npix, beta, ar, ai, m, t, tc);
  return hipGetLastError();
}

hipError_t launch_ilt_reg(const float* m, int n, int ntargets, int bpi, float tv_w, float tv_eps,
                          float bin_w, float* reg_grad, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_ilt_reg, dim3(bpi, ntargets), dim3(PW_THREADS), 0, s, m, n, bpi, tv_w,
                     tv_eps, bin_w, reg_grad, partial);
  return hipGetLastError();
}

hipError_t launch_ilt_cost_reduce(const double* fid, const double* reg, int ntargets, int ncond,
                                  int bpi, int reg_bpi, const float* weight, double inv_n,
                                  double* cost, hipStream_t s) {
  hipLaunchKernelGGL(k_ilt_cost_reduce, dim3((ntargets + 63) / 64), dim3(64), 0, s, fid, reg,
                     ntargets, ncond, bpi, reg_bpi, weight, inv_n, cost);
  return hipGetLastError();
}

template <int M>
static hipError_t adjoint_m(int nimg, int KS, const float2* Mq, const float2* Gb, int B, int ncond,
                            const int* cond_K, const float2* const* cond_kern,
                            const float* const* cond_lam, float lam_scale, float g0_scale,
                            float flare, float* G0g, float2* Rpart, hipStream_t s) {
  const int W = 2 * B + 1, W2 = 4 * B + 1;
  const size_t sh0 = (size_t)W2 * (M + 1) * sizeof(c32);
  const size_t sh1 = (size_t)2 * W * (M + 1) * sizeof(c32) + (size_t)M * (M + 1) * sizeof(float);
  auto k0 = k_ilt_g0<M>;
  auto k1 = k_ilt_adjoint<M>;
  // sized for the largest band (B = 15), once per instantiation
  static const hipError_t a0 = hipFuncSetAttribute(
      (const void*)k0, hipFuncAttributeMaxDynamicSharedMemorySize, (int)((61 * (M + 1)) * sizeof(c32)));
  static const hipError_t a1 = hipFuncSetAttribute(
      (const void*)k1, hipFuncAttributeMaxDynamicSharedMemorySize,
      (int)(2 * 31 * (M + 1) * sizeof(c32) + M * (M + 1) * sizeof(float)));
  if (a0 != hipSuccess) return a0;
  if (a1 != hipSuccess) return a1;
  hipLaunchKernelGGL(k0, dim3(nimg), dim3(M), sh0, s, Gb, B, g0_scale, flare, G0g);
  hipLaunchKernelGGL(k1, dim3(nimg, KS), dim3(M), sh1, s, Mq, G0g, B, ncond, cond_K, cond_kern,
                     cond_lam, lam_scale, Rpart);
  return hipGetLastError();
}

hipError_t launch_ilt_adjoint(int m, int nimg, int KS, const float2* Mq, const float2* Gb, int B,
                              int ncond, const int* cond_K, const float2* const* cond_kern,
                              const float* const* cond_lam, float lam_scale, float g0_scale,
                              float flare, float* G0g, float2* Rpart, hipStream_t s) {
  if (B > 15 || KS < 1) return hipErrorInvalidValue;
  switch (m) {
    case 32: return adjoint_m<32>(nimg, KS, Mq, Gb, B, ncond, cond_K, cond_kern, cond_lam, lam_scale, g0_scale, flare, G0g, Rpart, s);
    case 64: return adjoint_m<64>(nimg, KS, Mq, Gb, B, ncond, cond_K, cond_kern, cond_lam, lam_scale, g0_scale, flare, G0g, Rpart, s);
    default: return hipErrorInvalidValue;
  }
}

template <int M>
static hipError_t adjoint_big_m(int nimg, int KS, const float2* Mq, const float2* Gb, int B,
                                int ncond, const int* cond_K, const float2* const* cond_kern,
                                const float* const* cond_lam, float lam_scale, float g0_scale,
                                float flare, const float2* tw_m, float2* g0tmp, float* G0g,
                                float2* buf, float2* Rpart, hipStream_t s) {
  using F = AdjFft<M>;
  const int W = 2 * B + 1, W2 = 4 * B + 1, jobs = nimg * KS;
  auto blocks = [](int count, int teams) { return (unsigned)((count + teams - 1) / teams); };
  hipLaunchKernelGGL(k_ilt_g0_rows<M>, dim3(blocks(W2, F::TEAMS), nimg), dim3(256), 0, s, Gb, B,
                     flare, tw_m, g0tmp);
  hipLaunchKernelGGL(k_ilt_g0_cols<M>, dim3(blocks(M, F::TEAMS), nimg), dim3(256), 0, s, g0tmp, B,
                     g0_scale, tw_m, G0g);
  hipLaunchKernelGGL(k_ilt_adj_rows<M>, dim3(blocks(W, F::TEAMS), jobs), dim3(256), 0, s, Mq, B,
                     ncond, KS, cond_K, cond_kern, tw_m, buf);
  hipLaunchKernelGGL(k_ilt_adj_cols<M>, dim3(blocks(M, kAdjColTeams), jobs),
                     dim3(kAdjColTeams * F::N2), 0, s, G0g, B, KS, tw_m, buf);
  hipLaunchKernelGGL(k_ilt_adj_out<M>, dim3(blocks(W, F::TEAMS), jobs), dim3(256), 0, s, buf, B,
                     ncond, KS, cond_K, cond_kern, cond_lam, lam_scale, tw_m, Rpart);
  return hipGetLastError();
}

hipError_t launch_ilt_adjoint_big(int m, int nimg, int KS, const float2* Mq, const float2* Gb,
                                  int B, int ncond, const int* cond_K,
                                  const float2* const* cond_kern, const float* const* cond_lam,
                                  float lam_scale, float g0_scale, float flare,
                                  const float2* tw_m, float2* g0tmp, float* G0g, float2* buf,
                                  float2* Rpart, hipStream_t s) {
  if (KS < 1 || nimg < 1 || (int64_t)nimg * KS > 65535 || 4 * B + 1 > m) return hipErrorInvalidValue;
  switch (m) {
    case 128: return adjoint_big_m<128>(nimg, KS, Mq, Gb, B, ncond, cond_K, cond_kern, cond_lam, lam_scale, g0_scale, flare, tw_m, g0tmp, G0g, buf, Rpart, s);
    case 192: return adjoint_big_m<192>(nimg, KS, Mq, Gb, B, ncond, cond_K, cond_kern, cond_lam, lam_scale, g0_scale, flare, tw_m, g0tmp, G0g, buf, Rpart, s);
    case 256: return adjoint_big_m<256>(nimg, KS, Mq, Gb, B, ncond, cond_K, cond_kern, cond_lam, lam_scale, g0_scale, flare, tw_m, g0tmp, G0g, buf, Rpart, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_ilt_hband(const float2* Rin, int B, int ntargets, int ncond, int KS,
                            const float* weight, float cr, float ci, float inv_n2, float2* H,
                            hipStream_t s) {
  const int cnt = (B + 1) * (2 * B + 1);
  hipLaunchKernelGGL(k_ilt_hband, dim3((cnt + 127) / 128, ntargets), dim3(128), 0, s, Rin, B,
                     ncond, KS, weight, cr, ci, inv_n2, H);
  return hipGetLastError();
}

hipError_t launch_ilt_filter_band(const float2* X, int Bf, float c, float inv_n2, int count,
                                  float2* Ib, hipStream_t s) {
  const int64_t total = (int64_t)count * (Bf + 1) * (2 * Bf + 1);
  hipLaunchKernelGGL(k_ilt_filter_band, dim3(grid_for(total, 256)), dim3(256), 0, s, X, Bf, c,
                     inv_n2, count, Ib);
  return hipGetLastError();
}

hipError_t launch_ilt_proxy(const float* G, int ntargets, int ncond, int64_t npix, int bpi,
                            const float* weight, float flare, double* part, float* dcdm,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_ilt_gsum, dim3(bpi, ntargets * ncond), dim3(256), 0, s, G, npix, bpi,
                     part);
  hipLaunchKernelGGL(k_ilt_proxy, dim3(grid_for((int64_t)ntargets * npix, 256)), dim3(256), 0, s,
                     G, part, ntargets, ncond, npix, bpi, weight, flare, dcdm);
  return hipGetLastError();
}

static dim3 rows_grid(int64_t npix, int count) {
  const int64_t bx = std::min<int64_t>((npix + 255) / 256, 1024);
  return dim3((unsigned)bx, (unsigned)count);
}

hipError_t launch_ilt_chain(const float* dcdm, const float* reg, const float* m, float beta,
                            int64_t npix, int count, const int* orows, float* grad,
                            hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ilt_chain, rows_grid(npix, count), dim3(256), 0, s, dcdm, reg, m, beta,
                     npix, orows, grad);
  return hipGetLastError();
}

hipError_t launch_vec_axpy_rows(float* x, const float* d, const double* coef, int64_t npix,
                                int count, const int* rows, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_vec_axpy_rows, rows_grid(npix, count), dim3(256), 0, s, x, d, coef, npix,
                     rows);
  return hipGetLastError();
}

hipError_t launch_vec_reduce(const float* a, const float* b, int64_t npix, int count,
                             const int* rows, int nblk, int ops, int nops, int T, double* part,
                             double* out, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (nops < 1 || nops > 3 || count > T) return hipErrorInvalidValue;
  for (int r = 0; r < nops; ++r) {
    const int op = (ops >> (4 * r)) & 15;
    if (op > 3 || ((op == 1 || op == 2) && !b)) return hipErrorInvalidValue;
  }
  const dim3 grid(nblk, count);
  if (nops == 1)
    hipLaunchKernelGGL(k_vec_reduce<1>, grid, dim3(PW_THREADS), 0, s, a, b, npix, nblk, ops, T, rows, part);
  else if (nops == 2)
    hipLaunchKernelGGL(k_vec_reduce<2>, grid, dim3(PW_THREADS), 0, s, a, b, npix, nblk, ops, T, rows, part);
  else
    hipLaunchKernelGGL(k_vec_reduce<3>, grid, dim3(PW_THREADS), 0, s, a, b, npix, nblk, ops, T, rows, part);
  if (out)
    hipLaunchKernelGGL(k_vec_finish, dim3((count + 63) / 64, nops), dim3(64), 0, s, part, count,
                       nblk, ops, nops, T, rows, out);
  return hipGetLastError();
}

hipError_t launch_vec_dir(const float* g, const double* a, const double* b, int64_t npix,
                          int count, const int* rows, float* d, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_vec_dir, rows_grid(npix, count), dim3(256), 0, s, g, a, b, npix, rows, d);
  return hipGetLastError();
}

hipError_t launch_vec_dir_reduce(const float* g, const double* a, const double* b, int64_t npix,
                                 int count, const int* rows, int nblk, int T, float* d,
                                 double* part, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (count > T || nblk < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_vec_dir_reduce, dim3(nblk, count), dim3(PW_THREADS), 0, s, g, a, b, npix,
                     nblk, T, rows, d, part);
  return hipGetLastError();
}

hipError_t launch_rows_gather(const float* src, const int* idx, int64_t npix, int nrows,
                              float* dst, hipStream_t s) {
  hipLaunchKernelGGL(k_rows_move<false>, dim3(grid_for(npix * nrows, 256)), dim3(256), 0, s, src,
                     idx, npix, nrows, dst);
  return hipGetLastError();
}

hipError_t launch_rows_scatter(const float* src, const int* idx, int64_t npix, int nrows,
                               float* dst, hipStream_t s) {
  hipLaunchKernelGGL(k_rows_move<true>, dim3(grid_for(npix * nrows, 256)), dim3(256), 0, s, src,
                     idx, npix, nrows, dst);
  return hipGetLastError();
}

hipError_t launch_vec_axpy_gather(const float* x, const float* d, const double* coef, int64_t npix,
                                  int count, const int* rows, float* out, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_vec_axpy_gather, rows_grid(npix, count), dim3(256), 0, s, x, d, coef, npix,
                     rows, out);
  return hipGetLastError();
}

}  // namespace huvk
