// HIP kernels of the band-limited SOCS image path (gfx950).
//
// Reference semantics (crates/highuvlith-core/src/aerial.rs:1682-1759):
//   I = (1 - phi) * sum_k lambda_k |IFFT_n(K_k . M)|^2 / I_clear + phi * <.>
// with M = FFT_n(mask) (compute_from_transmittance, aerial.rs:1789-1799).
//
// Every kernel K_k is zero outside the disk |f| <= r, so each field E_k is a
// trigonometric polynomial of per-axis degree B = floor(r) and |E_k|^2 one of
// degree 2B.  Sampling on a coarse m x m grid with m >= 4B + 1 is therefore
// exact (no aliasing): the fields are synthesised and squared on that grid
// and never leave the chip; only the coarse intensity is transformed back
// once per image and evaluated on the n x n output grid.
//
// Stages (one launch each, all batched over clips / images):
//   rows_band    mask rows (real, FP32) -> row spectra, band |kx| <= B kept
//   cols_band    row spectra -> M[ky][kx] band, twisted per x class q
//   socs_fields  K kernels x band -> coarse intensity (fields in registers)
//   coarse_rows  coarse rows -> |gx| <= 2B
//   coarse_cols  -> intensity band I^[gy][gx], 1/m^2, flare on non-DC terms
//   band_image   band -> n x n image rows (C2R), written once to HBM
#include <hip/hip_runtime.h>

#include "fft_lds.hpp"
#include "socs_kernels.hpp"

namespace huvk {

// ---------------------------------------------------------------------------
// rows_band: two real rows packed as z = a + i b, FFT_n forward, keep bins
// k in [0, B] and [n - B, n).  tmp layout: [clip][kk][pair], kk in [0, 2B],
// kk <= B -> k = kk, kk > B -> k = n - (kk - B).
// ---------------------------------------------------------------------------
template <int N1, int N2, int TEAMS>
__global__ __launch_bounds__(TEAMS * N2) void k_rows_band(const float* __restrict__ masks, int n,
                                                          int B, const float2* __restrict__ tw_n,
                                                          float2* __restrict__ tmp) {
  constexpr int N = N1 * N2;
  __shared__ c32 scratch[TEAMS][(N1 + 1) * N2];
  const int team = threadIdx.x / N2, t = threadIdx.x % N2;
  const int clip = blockIdx.y;
  const int pair = blockIdx.x * TEAMS + team;
  const int npairs = N / 2;
  const bool active = pair < npairs;
  const float* ra = masks + ((size_t)clip * N + 2 * (size_t)(active ? pair : 0)) * N;
  const float* rb = ra + N;
  const int nk = 2 * B + 1;
  float2* out = tmp + (size_t)clip * nk * npairs;
  fft_four_step<-1, N1, N2>(
      t, scratch[team], tw_n,
      [&](int i) { return active ? c32{ra[i], rb[i]} : c32{0.f, 0.f}; },
      [&](int k, c32 v) {
        if (!active) return;
        int kk = -1;
        if (k <= B) kk = k;
        else if (k >= N - B) kk = B + (N - k);
        if (kk >= 0) out[(size_t)kk * npairs + pair] = make_float2(v.re, v.im);
      });
  (void)n;
}

// ---------------------------------------------------------------------------
// cols_band: for column kx in [0, B], assemble the column of row spectra
// from the packed pairs, FFT_n forward along y, keep ky in [-B, B].  Writes
// Mq[clip][q][kx + B][ky + B] = M[ky][kx] * pt * e^{+2 pi i kx q / m} and the
// Hermitian mirror at (-kx, -ky).  pt = pixel transfer sinc(pi kx/n) sinc(pi ky/n)
// (table sinc_t[k + B], all ones when disabled).
// ---------------------------------------------------------------------------
template <int N1, int N2, int TEAMS>
__global__ __launch_bounds__(TEAMS * N2) void k_cols_band(
    const float2* __restrict__ tmp, int B, int m, int Q, const float2* __restrict__ tw_n,
    const float2* __restrict__ tw_m, const float* __restrict__ sinc_t, float2* __restrict__ Mq) {
  constexpr int N = N1 * N2;
  __shared__ c32 scratch[TEAMS][(N1 + 1) * N2];
  const int team = threadIdx.x / N2, t = threadIdx.x % N2;
  const int clip = blockIdx.y;
  const int kx = blockIdx.x * TEAMS + team;
  const bool active = kx <= B;
  const int kxa = active ? kx : 0;
  const int npairs = N / 2;
  const int nk = 2 * B + 1;
  const int W = 2 * B + 1;
  const float2* col_pos = tmp + ((size_t)clip * nk + kxa) * npairs;
  const float2* col_neg = tmp + ((size_t)clip * nk + (kxa == 0 ? 0 : B + kxa)) * npairs;
  float2* mq = Mq + (size_t)clip * Q * W * W;
  fft_four_step<-1, N1, N2>(
      t, scratch[team], tw_n,
      [&](int i) {
        const int p = i >> 1;
        const float2 zp = col_pos[p], zn = col_neg[p];
        // A = (Z[k] + conj Z[n-k]) / 2 ; Bv = (Z[k] - conj Z[n-k]) / (2i)
        if ((i & 1) == 0) return c32{0.5f * (zp.x + zn.x), 0.5f * (zp.y - zn.y)};
        return c32{0.5f * (zp.y + zn.y), -0.5f * (zp.x - zn.x)};
      },
      [&](int k, c32 v) {
        if (!active) return;
        int ky;
        if (k <= B) ky = k;
        else if (k >= N - B) ky = k - N;
        else return;
        const float pt = sinc_t[kx + B] * sinc_t[ky + B];
        v = cscale(v, pt);
        for (int q = 0; q < Q; ++q) {
          const c32 w = tw_lookup<+1>(tw_m, (kx * q) % m);
          const c32 a = cmul(v, w);
          mq[((size_t)q * W + (kx + B)) * W + (ky + B)] = make_float2(a.re, a.im);
          if (kx > 0) {
            // conj(M) e^{-2 pi i kx q / m} = conj(M e^{+...})
            mq[((size_t)q * W + (B - kx)) * W + (B - ky)] = make_float2(a.re, -a.im);
          }
        }
      });
}

// ---------------------------------------------------------------------------
// socs_fields: one work-group per (clip, x class q); all planes of the clip.
// FX = M/Q columns per class, FY = M/R rows per y task.  Threads: FX * R.
// x task (band row iy):  c[h] = sum_{kx = h mod FX} K[k][kx][iy] Mq[q][kx][iy];
//                        T[iy][j] = IFFT_FX(c)[j]   (E row at x = Q j + q)
// y task (column j, residue r):
//                        d[h] = e^{2 pi i h r/M} sum_{ky = h mod FY} T[ky][j] e^{2 pi i t r/R}
//                        E[R j' + r][Q j + q] = IFFT_FY(d)[j'];  acc += lambda |E|^2
// Coarse intensity layout: coarse[img][q][y][j].
// ---------------------------------------------------------------------------
template <int M, int Q, int R>
struct FieldCfg {
  static constexpr int FX = M / Q;
  static constexpr int FY = M / R;
  static constexpr int THREADS = FX * R;
};

// d[h] *= e^{SIGN 2 pi i h r / M} for h < FY (compile-time r).
template <int SIGN, int M, int r, int H, int FY>
__device__ __forceinline__ void ytask_twiddle(c32* d) {
  if constexpr (H < FY) {
    d[H] = twiddle<SIGN, (long long)H * r, M>(d[H]);
    ytask_twiddle<SIGN, M, r, H + 1, FY>(d);
  }
}

// y task of socs_fields for residue r: fold the band rows of column j into
// FY bins, twist, inverse FFT_FY, accumulate lambda |E|^2.
template <int M, int R, int FY, int TP, int r>
__device__ __forceinline__ void ytask(const c32* __restrict__ T, int B, int j, float l,
                                      float* acc) {
  c32 d[FY];
#pragma unroll
  for (int h = 0; h < FY; ++h) {
    d[h] = c32{0.f, 0.f};
    if (h <= B) d[h] = T[(h + B) * TP + j];
  }
  // negative ky = h - FY carries e^{-2 pi i r / R}
#pragma unroll
  for (int h = 0; h < FY; ++h) {
    if (h - FY >= -B) d[h] = cadd(d[h], twiddle<+1, -(long long)r, R>(T[(h - FY + B) * TP + j]));
  }
  ytask_twiddle<+1, M, r, 0, FY>(d);
  fft_reg<+1, FY>(d);
#pragma unroll
  for (int i = 0; i < FY; ++i) acc[i] += l * cnorm(d[i]);
}

template <int M, int Q, int R>
__global__ __launch_bounds__((FieldCfg<M, Q, R>::THREADS)) void k_socs_fields(
    const float2* __restrict__ Mq, int B, int nplanes, const int* __restrict__ plane_K,
    const float2* const* __restrict__ plane_kern, const float* const* __restrict__ plane_lam,
    float* __restrict__ coarse) {
  using C = FieldCfg<M, Q, R>;
  constexpr int FX = C::FX, FY = C::FY;
  constexpr int TP = FX + 1;  // padded T row
  extern __shared__ c32 T[];  // (2B+1) x TP
  const int q = blockIdx.x, clip = blockIdx.y;
  const int W = 2 * B + 1;
  const int tid = threadIdx.x;
  const float2* mq = Mq + ((size_t)clip * Q + q) * W * W;
  const int j = tid % FX, r = tid / FX;

  for (int p = 0; p < nplanes; ++p) {
    const int K = plane_K[p];
    const float2* kern = plane_kern[p];
    const float* lam = plane_lam[p];
    float acc[FY];
#pragma unroll
    for (int i = 0; i < FY; ++i) acc[i] = 0.f;
    for (int k = 0; k < K; ++k) {
      // ---- x pass ----
      for (int iy = tid; iy < W; iy += C::THREADS) {
        const float2* kk = kern + (size_t)k * W * W + iy;
        const float2* mm = mq + iy;
        c32 c[FX];
#pragma unroll
        for (int h = 0; h < FX; ++h) {
          c[h] = c32{0.f, 0.f};
          // kx = h (t = 0) and kx = h - FX (t = -1)
          if (h <= B) {
            const float2 a = kk[(size_t)(h + B) * W], b = mm[(size_t)(h + B) * W];
            c[h] = cadd(c[h], cmul(c32{a.x, a.y}, c32{b.x, b.y}));
          }
          if (h - FX >= -B) {
            const float2 a = kk[(size_t)(h - FX + B) * W], b = mm[(size_t)(h - FX + B) * W];
            c[h] = cadd(c[h], cmul(c32{a.x, a.y}, c32{b.x, b.y}));
          }
        }
        fft_reg<+1, FX>(c);
#pragma unroll
        for (int jj = 0; jj < FX; ++jj) T[iy * TP + jj] = c[jj];
      }
      __syncthreads();
      // ---- y pass ----
      {
        const float l = lam[k];
        switch (r) {
          case 0: ytask<M, R, FY, TP, 0>(T, B, j, l, acc); break;
          case 1: if constexpr (R > 1) ytask<M, R, FY, TP, 1>(T, B, j, l, acc); break;
          case 2: if constexpr (R > 2) ytask<M, R, FY, TP, 2>(T, B, j, l, acc); break;
          case 3: if constexpr (R > 3) ytask<M, R, FY, TP, 3>(T, B, j, l, acc); break;
          default: break;
        }
      }
      __syncthreads();
    }
    float* out = coarse + (((size_t)clip * nplanes + p) * Q + q) * M * FX;
#pragma unroll
    for (int i = 0; i < FY; ++i) out[(size_t)(R * i + r) * FX + j] = acc[i];
  }
}

// ---------------------------------------------------------------------------
// coarse_rows: coarse row pairs -> FFT_M forward, keep gx in [0, 2B] and
// [M - 2B, M).  tmp1 layout [img][kk][pair] with kk in [0, 4B].
// ---------------------------------------------------------------------------
template <int M, int Q, int N1, int N2, int TEAMS>
__global__ __launch_bounds__((TEAMS * team_size<N1, N2>())) void k_coarse_rows(
    const float* __restrict__ coarse, int B, const float2* __restrict__ tw_m,
    float2* __restrict__ tmp1) {
  constexpr int TS = team_size<N1, N2>();
  constexpr int FX = M / Q;
  __shared__ c32 scratch[TEAMS][(N1 + 1) * N2];
  const int team = threadIdx.x / TS, t = threadIdx.x % TS;
  const int img = blockIdx.y;
  const int pair = blockIdx.x * TEAMS + team;
  const int npairs = M / 2;
  const bool active = pair < npairs;
  const int ya = 2 * (active ? pair : 0), yb = ya + 1;
  const float* cimg = coarse + (size_t)img * M * M;
  const int G = 2 * B;
  const int nk = 2 * G + 1;
  float2* out = tmp1 + (size_t)img * nk * npairs;
  fft_four_step<-1, N1, N2>(
      t, scratch[team], tw_m,
      [&](int x) {
        if (!active) return c32{0.f, 0.f};
        const int qq = x % Q, jj = x / Q;
        const float* base = cimg + (size_t)qq * M * FX;
        return c32{base[(size_t)ya * FX + jj], base[(size_t)yb * FX + jj]};
      },
      [&](int k, c32 v) {
        if (!active) return;
        int kk = -1;
        if (k <= G) kk = k;
        else if (k >= M - G) kk = G + (M - k);
        if (kk >= 0) out[(size_t)kk * npairs + pair] = make_float2(v.re, v.im);
      });
}

// coarse_cols: column gx in [0, 2B]: separate the packed pairs, FFT_M along y,
// keep gy in [-2B, 2B]; Ib[img][gx][gy + 2B] = v / M^2, times (1 - phi) except DC.
template <int M, int N1, int N2, int TEAMS>
__global__ __launch_bounds__((TEAMS * team_size<N1, N2>())) void k_coarse_cols(
    const float2* __restrict__ tmp1, int B, float flare, const float2* __restrict__ tw_m,
    float2* __restrict__ Ib) {
  constexpr int TS = team_size<N1, N2>();
  __shared__ c32 scratch[TEAMS][(N1 + 1) * N2];
  const int team = threadIdx.x / TS, t = threadIdx.x % TS;
  const int img = blockIdx.y;
  const int G = 2 * B;
  const int gx = blockIdx.x * TEAMS + team;
  const bool active = gx <= G;
  const int gxa = active ? gx : 0;
  const int npairs = M / 2;
  const int nk = 2 * G + 1;
  const float2* col_pos = tmp1 + ((size_t)img * nk + gxa) * npairs;
  const float2* col_neg = tmp1 + ((size_t)img * nk + (gxa == 0 ? 0 : G + gxa)) * npairs;
  float2* out = Ib + ((size_t)img * (G + 1) + gxa) * nk;
  const float inv = 1.0f / ((float)M * (float)M);
  fft_four_step<-1, N1, N2>(
      t, scratch[team], tw_m,
      [&](int i) {
        const int p = i >> 1;
        const float2 zp = col_pos[p], zn = col_neg[p];
        if ((i & 1) == 0) return c32{0.5f * (zp.x + zn.x), 0.5f * (zp.y - zn.y)};
        return c32{0.5f * (zp.y + zn.y), -0.5f * (zp.x - zn.x)};
      },
      [&](int k, c32 v) {
        if (!active) return;
        int gy;
        if (k <= G) gy = k;
        else if (k >= M - G) gy = k - M;
        else return;
        const float s = (gx == 0 && gy == 0) ? inv : inv * (1.0f - flare);
        out[gy + G] = make_float2(v.re * s, v.im * s);
      });
}

// ---------------------------------------------------------------------------
// band_image: rows y = P j + p (P = n / 64) of the n x n image.
// Phase 1 (threads gx in [0, G]):  U[j][gx] = IFFT_64( fold_64( Ib[gx][gy] e^{2 pi i gy p / n} ) )
// Phase 2: row pairs via C2R packing, four-step IFFT_n, coalesced row writes.
// ---------------------------------------------------------------------------
template <int N1, int N2, int TEAMS>
__global__ __launch_bounds__(256) void k_band_image(const float2* __restrict__ Ib, int n, int G,
                                                    const float2* __restrict__ tw_n,
                                                    float* __restrict__ out) {
  constexpr int N = N1 * N2;
  constexpr int JR = 64;  // rows per work-group
  extern __shared__ c32 smem[];
  // G: half-width of the (Hermitian) intensity band
  const int nk = 2 * G + 1;
  const int UP = G + 2;  // padded U row
  c32* U = smem;                                  // JR x UP
  c32* wt = smem + JR * UP;                       // nk twiddles
  c32* scratch = wt + ((nk + 1) & ~1);            // TEAMS x (N1+1) N2
  const int p = blockIdx.x, img = blockIdx.y;
  const int P = N / JR;
  const int tid = threadIdx.x;
  for (int i = tid; i < nk; i += 256) {
    const int gy = i - G;
    const int e = (int)(((long long)gy * p % N + N) % N);
    wt[i] = tw_lookup<+1>(tw_n, e);
  }
  __syncthreads();
  const float2* ib = Ib + (size_t)img * (G + 1) * nk;
  for (int gx = tid; gx <= G; gx += 256) {
    const float2* col = ib + (size_t)gx * nk;
    c32 d[JR];
#pragma unroll
    for (int h = 0; h < JR; ++h) {
      d[h] = c32{0.f, 0.f};
#pragma unroll
      for (int tt = -2; tt <= 1; ++tt) {
        const int gy = h + JR * tt;
        if (gy >= -G && gy <= G) {
          const float2 v = col[gy + G];
          d[h] = cadd(d[h], cmul(c32{v.x, v.y}, wt[gy + G]));
        }
      }
    }
    fft_reg<+1, JR>(d);
#pragma unroll
    for (int jj = 0; jj < JR; ++jj) U[jj * UP + gx] = d[jj];
  }
  __syncthreads();
  constexpr int TS = team_size<N1, N2>();
  const int team = tid / TS, t = tid % TS;
  float* img_out = out + (size_t)img * N * N;
  for (int base = 0; base < JR; base += 2 * TEAMS) {
    const int ja = base + 2 * team;
    const c32* ua = U + ja * UP;
    const c32* ub = ua + UP;
    float* ra = img_out + (size_t)(P * ja + p) * N;
    float* rb = img_out + (size_t)(P * (ja + 1) + p) * N;
    fft_four_step<+1, N1, N2>(
        t, scratch + team * (N1 + 1) * N2, tw_n,
        [&](int i) {
          const int g = i < N / 2 ? i : i - N;
          if (g > G || g < -G) return c32{0.f, 0.f};
          if (g >= 0) {
            const c32 a = ua[g], b = ub[g];
            return c32{a.re - b.im, a.im + b.re};  // a + i b
          }
          const c32 a = ua[-g], b = ub[-g];
          return c32{a.re + b.im, -a.im + b.re};  // conj(a) + i conj(b)
        },
        [&](int k, c32 v) {
          ra[k] = v.re;
          rb[k] = v.im;
        });
  }
}

}  // namespace huvk

// ===========================================================================
// Launchers
// ===========================================================================
namespace huvk {

int choose_coarse(int n, int B) {
  const int sizes[] = {32, 64, 128, 192, 256};
  for (int m : sizes)
    if (m >= 4 * B + 1 && m <= n && B < m / coarse_classes(m)) return m;
  return 0;
}

int coarse_classes(int m) {
  switch (m) {
    case 32: case 64: return 1;
    case 128: return 2;
    case 192: return 3;
    case 256: return 4;
    default: return 0;
  }
}

#define HUV_LAUNCH(kern, grid, block, shmem, s, ...)                   \
  do {                                                                 \
    hipLaunchKernelGGL(kern, grid, block, shmem, s, __VA_ARGS__);      \
    return hipGetLastError();                                          \
  } while (0)

template <int N1, int N2>
static hipError_t rows_band_n(int B, int nclips, const float* masks, const float2* tw,
                              float2* tmp, hipStream_t s) {
  constexpr int TEAMS = 128 / N2 > 0 ? 128 / N2 : 1;
  const int npairs = N1 * N2 / 2;
  dim3 grid((npairs + TEAMS - 1) / TEAMS, nclips);
  HUV_LAUNCH((k_rows_band<N1, N2, TEAMS>), grid, dim3(TEAMS * N2), 0, s, masks, N1 * N2, B, tw,
             tmp);
}

hipError_t launch_rows_band(int n, int B, int nclips, const float* masks, const float2* tw_n,
                            float2* tmp, hipStream_t s) {
  switch (n) {
    case 256: return rows_band_n<16, 16>(B, nclips, masks, tw_n, tmp, s);
    case 512: return rows_band_n<16, 32>(B, nclips, masks, tw_n, tmp, s);
    case 1024: return rows_band_n<32, 32>(B, nclips, masks, tw_n, tmp, s);
    case 2048: return rows_band_n<32, 64>(B, nclips, masks, tw_n, tmp, s);
    default: return hipErrorInvalidValue;
  }
}

template <int N1, int N2>
static hipError_t cols_band_n(int B, int m, int nclips, const float2* tmp, const float2* tw_n,
                              const float2* tw_m, const float* sinc_t, float2* Mq,
                              hipStream_t s) {
  constexpr int TEAMS = 128 / N2 > 0 ? 128 / N2 : 1;
  dim3 grid((B + 1 + TEAMS - 1) / TEAMS, nclips);
  HUV_LAUNCH((k_cols_band<N1, N2, TEAMS>), grid, dim3(TEAMS * N2), 0, s, tmp, B, m,
             coarse_classes(m), tw_n, tw_m, sinc_t, Mq);
}

hipError_t launch_cols_band(int n, int B, int m, int nclips, const float2* tmp,
                            const float2* tw_n, const float2* tw_m, const float* sinc_t,
                            float2* Mq, hipStream_t s) {
  switch (n) {
    case 256: return cols_band_n<16, 16>(B, m, nclips, tmp, tw_n, tw_m, sinc_t, Mq, s);
    case 512: return cols_band_n<16, 32>(B, m, nclips, tmp, tw_n, tw_m, sinc_t, Mq, s);
    case 1024: return cols_band_n<32, 32>(B, m, nclips, tmp, tw_n, tw_m, sinc_t, Mq, s);
    case 2048: return cols_band_n<32, 64>(B, m, nclips, tmp, tw_n, tw_m, sinc_t, Mq, s);
    default: return hipErrorInvalidValue;
  }
}

template <int M, int Q, int R>
static hipError_t fields_m(int B, int nclips, int nplanes, const int* plane_K,
                           const float2* const* plane_kern, const float* const* plane_lam,
                           const float2* Mq, float* coarse, hipStream_t s) {
  using C = FieldCfg<M, Q, R>;
  const size_t shmem = (size_t)(2 * B + 1) * (C::FX + 1) * sizeof(c32);
  auto kern = k_socs_fields<M, Q, R>;
  hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
  HUV_LAUNCH(kern, dim3(Q, nclips), dim3(C::THREADS), shmem, s, Mq, B, nplanes, plane_K,
             plane_kern, plane_lam, coarse);
}

hipError_t launch_socs_fields(int m, int B, int nclips, int nplanes, const int* plane_K,
                              const float2* const* plane_kern, const float* const* plane_lam,
                              const float2* Mq, float* coarse, hipStream_t s) {
  switch (m) {
    case 32: return fields_m<32, 1, 1>(B, nclips, nplanes, plane_K, plane_kern, plane_lam, Mq, coarse, s);
    case 64: return fields_m<64, 1, 1>(B, nclips, nplanes, plane_K, plane_kern, plane_lam, Mq, coarse, s);
    case 128: return fields_m<128, 2, 2>(B, nclips, nplanes, plane_K, plane_kern, plane_lam, Mq, coarse, s);
    case 192: return fields_m<192, 3, 3>(B, nclips, nplanes, plane_K, plane_kern, plane_lam, Mq, coarse, s);
    case 256: return fields_m<256, 4, 4>(B, nclips, nplanes, plane_K, plane_kern, plane_lam, Mq, coarse, s);
    default: return hipErrorInvalidValue;
  }
}

template <int M, int Q, int N1, int N2>
static hipError_t coarse_rows_m(int B, int nimg, const float* coarse, const float2* tw_m,
                                float2* tmp1, hipStream_t s) {
  constexpr int TS = team_size<N1, N2>();
  constexpr int TEAMS = TS == 1 ? 64 : 256 / TS;
  dim3 grid((M / 2 + TEAMS - 1) / TEAMS, nimg);
  HUV_LAUNCH((k_coarse_rows<M, Q, N1, N2, TEAMS>), grid, dim3(TEAMS * TS), 0, s, coarse, B, tw_m,
             tmp1);
}

hipError_t launch_coarse_rows(int m, int B, int nimg, const float* coarse, const float2* tw_m,
                              float2* tmp1, hipStream_t s) {
  switch (m) {
    case 32: return coarse_rows_m<32, 1, 1, 32>(B, nimg, coarse, tw_m, tmp1, s);
    case 64: return coarse_rows_m<64, 1, 1, 64>(B, nimg, coarse, tw_m, tmp1, s);
    case 128: return coarse_rows_m<128, 2, 2, 64>(B, nimg, coarse, tw_m, tmp1, s);
    case 192: return coarse_rows_m<192, 3, 3, 64>(B, nimg, coarse, tw_m, tmp1, s);
    case 256: return coarse_rows_m<256, 4, 4, 64>(B, nimg, coarse, tw_m, tmp1, s);
    default: return hipErrorInvalidValue;
  }
}

template <int M, int N1, int N2>
static hipError_t coarse_cols_m(int B, int nimg, float flare, const float2* tmp1,
                                const float2* tw_m, float2* Ib, hipStream_t s) {
  constexpr int TS = team_size<N1, N2>();
  constexpr int TEAMS = TS == 1 ? 64 : 256 / TS;
  dim3 grid((2 * B + 1 + TEAMS - 1) / TEAMS, nimg);
  HUV_LAUNCH((k_coarse_cols<M, N1, N2, TEAMS>), grid, dim3(TEAMS * TS), 0, s, tmp1, B, flare,
             tw_m, Ib);
}

hipError_t launch_coarse_cols(int m, int B, int nimg, float flare, const float2* tmp1,
                              const float2* tw_m, float2* Ib, hipStream_t s) {
  switch (m) {
    case 32: return coarse_cols_m<32, 1, 32>(B, nimg, flare, tmp1, tw_m, Ib, s);
    case 64: return coarse_cols_m<64, 1, 64>(B, nimg, flare, tmp1, tw_m, Ib, s);
    case 128: return coarse_cols_m<128, 2, 64>(B, nimg, flare, tmp1, tw_m, Ib, s);
    case 192: return coarse_cols_m<192, 3, 64>(B, nimg, flare, tmp1, tw_m, Ib, s);
    case 256: return coarse_cols_m<256, 4, 64>(B, nimg, flare, tmp1, tw_m, Ib, s);
    default: return hipErrorInvalidValue;
  }
}

template <int N1, int N2>
static hipError_t band_image_n(int G, int nimg, const float2* Ib, const float2* tw_n, float* out,
                               hipStream_t s) {
  constexpr int N = N1 * N2;
  constexpr int TS = team_size<N1, N2>();
  constexpr int TEAMS = 256 / TS;
  const int nk = 2 * G + 1;
  const size_t words = (size_t)64 * (G + 2) + ((nk + 1) & ~1) + (size_t)TEAMS * (N1 + 1) * N2;
  const size_t shmem = words * sizeof(c32);
  auto kern = k_band_image<N1, N2, TEAMS>;
  hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
  HUV_LAUNCH(kern, dim3(N / 64, nimg), dim3(256), shmem, s, Ib, N, G, tw_n, out);
}

hipError_t launch_band_image(int n, int G, int nimg, const float2* Ib, const float2* tw_n,
                             float* out, hipStream_t s) {
  switch (n) {
    case 256: return band_image_n<16, 16>(G, nimg, Ib, tw_n, out, s);
    case 512: return band_image_n<16, 32>(G, nimg, Ib, tw_n, out, s);
    case 1024: return band_image_n<32, 32>(G, nimg, Ib, tw_n, out, s);
    case 2048: return band_image_n<32, 64>(G, nimg, Ib, tw_n, out, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace huvk
