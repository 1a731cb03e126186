// Exact mask spectrum of rectangle masks, evaluated on the device.
//
// The spectrum convention is the one of mask.rs:31-45 (crates/highuvlith-core):
// for the field [-L/2, L/2)^2, L = N p, and the signed order m,
//   S[m] = N^2 c(m) e^{-i pi (mx + my)(N - 1)/N},
//   c(m) = (1/L^2) \iint t(x, y) e^{-2 pi i (mx x + my y)/L} dx dy,
// which equals FFT_N(raster) in the fine-pixel limit.  A rectangle cell
// [x0, x1) x [y0, y1) of weight w contributes w X(mx) Y(my) with
//   X(m) = N (wx/L) sinc(pi m wx/L) e^{-i pi m ((x0 + x1)/L + (N - 1)/N)}
// (mask.rs:1426-1477, axis_coefficients for a finite interval), and the
// background b adds b N^2 at m = 0 (mask.rs:633).  Axis tables are computed
// once per cell in FP64; the band box sums w X Y over the cells of a clip in
// cell order (FP64 accumulation, one FP32 rounding at the store).
#include <hip/hip_runtime.h>

#include "mask_kernels.hpp"

namespace huvk {

// N (w/L) sinc(pi m w/L) e^{-i pi m ((a + b)/L + (N - 1)/N)}
__device__ __forceinline__ double2 axis_coeff(double a, double b, int m, int n, double L) {
  const double w = b - a;
  const double u = (double)m * w / L;
  const double sinc = m == 0 ? 1.0 : sinpi(u) / (M_PI * u);
  const double amp = (double)n * (w / L) * sinc;
  // phase / (-pi), reduced to [-1, 1] before the sine and cosine
  double t = (double)m * ((a + b) / L + (double)(n - 1) / (double)n);
  t -= 2.0 * rint(0.5 * t);
  double sv, cv;
  sincospi(-t, &sv, &cv);
  return make_double2(amp * cv, amp * sv);
}

__global__ __launch_bounds__(256) void k_rect_tables(const double* __restrict__ cells,
                                                     int64_t ncells, int n, double L, int B,
                                                     double2* __restrict__ X,
                                                     double2* __restrict__ Y) {
  const int nk = 2 * B + 1;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncells * nk) return;
  const int64_t c = idx / nk;
  const int m = (int)(idx - c * nk) - B;
  const double* p = cells + 6 * c;
  const double2 xa = axis_coeff(p[0], p[1], m, n, L);
  const double2 ya = axis_coeff(p[2], p[3], m, n, L);
  const double wr = p[4], wi = p[5];
  X[idx] = make_double2(wr * xa.x - wi * xa.y, wr * xa.y + wi * xa.x);
  Y[idx] = ya;
}

// One thread per tile of kRectT consecutive rows (ky + BB) x kRectT consecutive
// columns (kx + BB), row tiles fastest: per cell it loads kRectT row and
// kRectT column coefficients for kRectT^2 products.  Work-groups
// [clip * bpc, (clip + 1) * bpc) cover one clip's box.
constexpr int kRectT = 4;

__global__ __launch_bounds__(256) void k_rect_band(const double2* __restrict__ X,
                                                   const double2* __restrict__ Y,
                                                   const int64_t* __restrict__ coff,
                                                   const double* __restrict__ bg, int n, int B,
                                                   int BB, int bpc, float2* __restrict__ Mq,
                                                   float2* __restrict__ Mp) {
  const int W = 2 * BB + 1, nk = 2 * B + 1;
  const int G = (W + kRectT - 1) / kRectT;
  const int clip = blockIdx.x / bpc;
  const int t = (blockIdx.x - clip * bpc) * blockDim.x + threadIdx.x;
  if (t >= G * G) return;
  const int gc = t / G, gr = t - gc * G;
  int xi[kRectT], yi[kRectT];
  bool xl[kRectT], yl[kRectT];
#pragma unroll
  for (int j = 0; j < kRectT; ++j) {
    const int kx = kRectT * gc + j - BB, ky = kRectT * gr + j - BB;
    xl[j] = kx >= -B && kx <= B && kRectT * gc + j < W;
    yl[j] = ky >= -B && ky <= B && kRectT * gr + j < W;
    // clamped: dead rows and columns compute and are discarded
    xi[j] = min(max(kx + B, 0), nk - 1);
    yi[j] = min(max(ky + B, 0), nk - 1);
  }
  double re[kRectT][kRectT], im[kRectT][kRectT];
#pragma unroll
  for (int i = 0; i < kRectT; ++i)
#pragma unroll
    for (int j = 0; j < kRectT; ++j) re[i][j] = im[i][j] = 0.0;
  const int64_t c0 = coff[clip], c1 = coff[clip + 1];
  const double2* xr = X + c0 * nk;
  const double2* yr = Y + c0 * nk;
  for (int64_t c = c0; c < c1; ++c, xr += nk, yr += nk) {
    double2 x[kRectT], y[kRectT];
#pragma unroll
    for (int j = 0; j < kRectT; ++j) {
      x[j] = xr[xi[j]];
      y[j] = yr[yi[j]];
    }
#pragma unroll
    for (int i = 0; i < kRectT; ++i)
#pragma unroll
      for (int j = 0; j < kRectT; ++j) {
        re[i][j] = fma(x[j].x, y[i].x, re[i][j]);
        re[i][j] = fma(-x[j].y, y[i].y, re[i][j]);
        im[i][j] = fma(x[j].x, y[i].y, im[i][j]);
        im[i][j] = fma(x[j].y, y[i].x, im[i][j]);
      }
  }
  const double nn = (double)n * (double)n;
  float2* mq = Mq + (size_t)clip * W * W;
  float2* mp = Mp ? Mp + (size_t)clip * (W + 1) * W : nullptr;
#pragma unroll
  for (int j = 0; j < kRectT; ++j) {
    const int col = kRectT * gc + j;
    if (col >= W) break;
#pragma unroll
    for (int i = 0; i < kRectT; ++i) {
      const int row = kRectT * gr + i;
      if (row >= W) break;
      const bool live = xl[j] && yl[i];
      double vr = live ? re[i][j] : 0.0, vi = live ? im[i][j] : 0.0;
      if (col == BB && row == BB) {  // kx = ky = 0: the background
        vr += bg[2 * clip] * nn;
        vi += bg[2 * clip + 1] * nn;
      }
      const float2 v = make_float2((float)vr, (float)vi);
      mq[(size_t)col * W + row] = v;
      if (mp) {
        // column-paired copy: (col, row) at [(col + 1) / 2][row][(col + 1) % 2], column -1 zero
        mp[((size_t)((col + 1) >> 1) * W + row) * 2 + ((col + 1) & 1)] = v;
        if (col == 0) mp[(size_t)row * 2] = make_float2(0.f, 0.f);
      }
    }
  }
}

// Simple polygons (mask.rs:1286-1352): for k = 2 pi m / L != 0
//   \iint_P e^{-ik.r} = (i/|k|^2) sum_e s (kx dy - ky dx) sinc(k.d/2) e^{-ik.r_mid}
// with s the polygon's orientation sign, and the area at k = 0.  edges:
// (dx, dy, cx, cy, re, im) with (re, im) = s (value - background); pdc per
// clip = sum of (value - background) area.  The box bins |kx|, |ky| <= B get
//   (N^2/L^2) e^{-i pi (mx + my)(N - 1)/N} x integral
// added to what the rectangle pass stored (plain and column-paired).
// One thread per tile of kRectT rows x kRectT columns: per edge it evaluates
// e^{-i kx cx} and the angle kx dx / 2 at its first column and steps them
// along the tile by angle addition (likewise along the rows), so each edge
// costs four sine/cosine pairs for kRectT^2 bins; sin(k.d/2) is the sine of
// the sum of the two half-angles, and sinc switches to its Taylor series
// below |u| = 1e-4, where that sum would lose relative accuracy.
__device__ __forceinline__ void cexp_neg(double t, double& c, double& s) {
  double sv, cv;
  sincos(t, &sv, &cv);
  c = cv;
  s = -sv;
}

__global__ __launch_bounds__(256) void k_poly_band(const double* __restrict__ edges,
                                                   const int64_t* __restrict__ eoff,
                                                   const double* __restrict__ pdc, int n,
                                                   double L, int B, int BB, int bpc,
                                                   float2* __restrict__ Mq,
                                                   float2* __restrict__ Mp) {
  const int W = 2 * BB + 1;
  const int G = (W + kRectT - 1) / kRectT;
  const int clip = blockIdx.x / bpc;
  const int t = (blockIdx.x - clip * bpc) * blockDim.x + threadIdx.x;
  if (t >= G * G) return;
  const int64_t e0 = eoff[clip], e1 = eoff[clip + 1];
  if (e0 == e1) return;
  const int gc = t / G, gr = t - gc * G;
  const int mx0 = kRectT * gc - BB, my0 = kRectT * gr - BB;
  const double dk = 2.0 * M_PI / L;
  const double kx0 = dk * mx0, ky0 = dk * my0;
  double ar[kRectT][kRectT], ai[kRectT][kRectT];
#pragma unroll
  for (int i = 0; i < kRectT; ++i)
#pragma unroll
    for (int j = 0; j < kRectT; ++j) ar[i][j] = ai[i][j] = 0.0;
  for (int64_t e = e0; e < e1; ++e) {
    const double* p = edges + 6 * e;
    const double dx = p[0], dy = p[1], cx = p[2], cy = p[3], wr = p[4], wi = p[5];
    // columns: w e^{-i kx cx} and (cos, sin)(kx dx / 2)
    double pxr[kRectT], pxi[kRectT], cax[kRectT], sax[kRectT];
    double pyr[kRectT], pyi[kRectT], cby[kRectT], sby[kRectT];
    {
      double c0, s0, cs, ss;
      cexp_neg(kx0 * cx, c0, s0);
      cexp_neg(dk * cx, cs, ss);
      pxr[0] = wr * c0 - wi * s0;
      pxi[0] = wr * s0 + wi * c0;
      double sa, ca, sd, cd;
      sincos(0.5 * kx0 * dx, &sa, &ca);
      sincos(0.5 * dk * dx, &sd, &cd);
      cax[0] = ca;
      sax[0] = sa;
#pragma unroll
      for (int j = 1; j < kRectT; ++j) {
        pxr[j] = pxr[j - 1] * cs - pxi[j - 1] * ss;
        pxi[j] = pxr[j - 1] * ss + pxi[j - 1] * cs;
        cax[j] = cax[j - 1] * cd - sax[j - 1] * sd;
        sax[j] = sax[j - 1] * cd + cax[j - 1] * sd;
      }
    }
    {
      double c0, s0, cs, ss;
      cexp_neg(ky0 * cy, c0, s0);
      cexp_neg(dk * cy, cs, ss);
      pyr[0] = c0;
      pyi[0] = s0;
      double sb, cb, sd, cd;
      sincos(0.5 * ky0 * dy, &sb, &cb);
      sincos(0.5 * dk * dy, &sd, &cd);
      cby[0] = cb;
      sby[0] = sb;
#pragma unroll
      for (int i = 1; i < kRectT; ++i) {
        pyr[i] = pyr[i - 1] * cs - pyi[i - 1] * ss;
        pyi[i] = pyr[i - 1] * ss + pyi[i - 1] * cs;
        cby[i] = cby[i - 1] * cd - sby[i - 1] * sd;
        sby[i] = sby[i - 1] * cd + cby[i - 1] * sd;
      }
    }
#pragma unroll
    for (int i = 0; i < kRectT; ++i) {
      const double ky = ky0 + dk * i;
#pragma unroll
      for (int j = 0; j < kRectT; ++j) {
        const double kx = kx0 + dk * j;
        const double kn = kx * dy - ky * dx;
        const double u = 0.5 * (kx * dx + ky * dy);
        const double su = sax[j] * cby[i] + cax[j] * sby[i];
        // su / u by the hardware reciprocal refined with two Newton steps
        double r = __builtin_amdgcn_rcp(u);
        r = fma(r, fma(-u, r, 1.0), r);
        r = fma(r, fma(-u, r, 1.0), r);
        const double sc = fabs(u) < 1e-4 ? 1.0 - u * u * (1.0 / 6.0) : su * r;
        const double a = kn * sc;
        const double hr = pxr[j] * pyr[i] - pxi[j] * pyi[i];
        const double hi = pxr[j] * pyi[i] + pxi[j] * pyr[i];
        ar[i][j] = fma(a, hr, ar[i][j]);
        ai[i][j] = fma(a, hi, ai[i][j]);
      }
    }
  }
  const double nf = (double)n;
  const double scale = nf * nf / (L * L);
  float2* mq = Mq + (size_t)clip * W * W;
  float2* mp = Mp ? Mp + (size_t)clip * (W + 1) * W : nullptr;
#pragma unroll
  for (int j = 0; j < kRectT; ++j) {
    const int col = kRectT * gc + j, mx = mx0 + j;
    if (col >= W) break;
#pragma unroll
    for (int i = 0; i < kRectT; ++i) {
      const int row = kRectT * gr + i, my = my0 + i;
      if (row >= W) break;
      if (mx < -B || mx > B || my < -B || my > B) continue;
      double fr, fi;
      if (mx == 0 && my == 0) {
        fr = pdc[2 * clip];
        fi = pdc[2 * clip + 1];
      } else {
        const double kx = dk * mx, ky = dk * my;
        const double inv = 1.0 / (kx * kx + ky * ky);  // i / |k|^2
        fr = -ai[i][j] * inv;
        fi = ar[i][j] * inv;
      }
      double tt = (double)(mx + my) * (nf - 1.0) / nf;
      tt -= 2.0 * rint(0.5 * tt);
      double sv, cv;
      sincospi(-tt, &sv, &cv);
      const double vr = scale * (fr * cv - fi * sv), vi = scale * (fr * sv + fi * cv);
      float2* q = mq + (size_t)col * W + row;
      const float2 old = *q;
      const float2 v = make_float2((float)((double)old.x + vr), (float)((double)old.y + vi));
      *q = v;
      if (mp) mp[((size_t)((col + 1) >> 1) * W + row) * 2 + ((col + 1) & 1)] = v;
    }
  }
}

hipError_t launch_poly_band(const double* edges, const int64_t* eoff, const double* pdc,
                            int nclips, int n, double L, int B, int BB, float2* Mq, float2* Mp,
                            hipStream_t s) {
  if (nclips <= 0) return hipSuccess;
  if (B < 0 || BB < B || BB > 1024) return hipErrorInvalidValue;
  const int W = 2 * BB + 1;
  const int G = (W + kRectT - 1) / kRectT;
  const int bpc = (G * G + 255) / 256;
  if ((int64_t)bpc * nclips >= (1ll << 31)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_poly_band, dim3((unsigned)(bpc * nclips)), dim3(256), 0, s, edges, eoff,
                     pdc, n, L, B, BB, bpc, Mq, Mp);
  return hipGetLastError();
}

hipError_t launch_rect_tables(const double* cells, int64_t ncells, int n, double L, int B,
                              double2* X, double2* Y, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  const int64_t total = ncells * (2 * B + 1);
  const int64_t blocks = (total + 255) / 256;
  if (blocks >= (1ll << 31)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_rect_tables, dim3((unsigned)blocks), dim3(256), 0, s, cells, ncells, n, L,
                     B, X, Y);
  return hipGetLastError();
}

hipError_t launch_rect_band(const double2* X, const double2* Y, const int64_t* coff,
                            const double* bg, int nclips, int n, int B, int BB, float2* Mq,
                            float2* Mp, hipStream_t s) {
  if (nclips <= 0) return hipSuccess;
  if (B < 0 || BB < B || BB > 1024) return hipErrorInvalidValue;
  const int W = 2 * BB + 1;
  const int G = (W + kRectT - 1) / kRectT;
  const int bpc = (G * G + 255) / 256;
  if ((int64_t)bpc * nclips >= (1ll << 31)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_rect_band, dim3((unsigned)(bpc * nclips)), dim3(256), 0, s, X, Y, coff,
                     bg, n, B, BB, bpc, Mq, Mp);
  return hipGetLastError();
}

}  // namespace huvk
