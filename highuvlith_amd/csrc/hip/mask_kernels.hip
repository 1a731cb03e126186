// Exact mask spectrum of masks given as rectangles and simple polygons,
// evaluated on the device (polygons: see k_poly_band below).
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

#include <algorithm>

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
// wpc work-groups per clip (of up to 512 threads), one thread per tile of
// kRectT rows x kRectT columns.  The clip's edges
// are taken in chunks; for each chunk the work-group first tabulates, per edge
// and order m in [-B, B], w e^{-i kx cx} with (cos, sin)(kx dx / 2) and
// e^{-i ky cy} with (cos, sin)(ky dy / 2) in LDS, then every tile reads its
// kRectT columns and rows from there: sin(k.d/2) is the sine of the sum of
// the two half-angles, sinc switches to its Taylor series below |u| = 1e-4
// (where that sum loses relative accuracy), and 1/u is the hardware
// reciprocal refined by two Newton steps.
__global__ __launch_bounds__(512) void k_poly_band(const double* __restrict__ edges,
                                                   const int64_t* __restrict__ eoff,
                                                   const double* __restrict__ pdc, int n,
                                                   double L, int B, int BB, int chunk, int wpc,
                                                   float2* __restrict__ Mq,
                                                   float2* __restrict__ Mp) {
  extern __shared__ double4 tab[];  // [axis][edge in chunk][m + B]
  const int W = 2 * BB + 1, nk = 2 * B + 1;
  const int G = (W + kRectT - 1) / kRectT;
  const int clip = blockIdx.x / wpc;
  const int t = (blockIdx.x - clip * wpc) * blockDim.x + threadIdx.x;  // tile index
  const int64_t e0 = eoff[clip], e1 = eoff[clip + 1];
  if (e0 == e1) return;  // uniform over the work-group
  const bool active = t < G * G;
  const int gc = active ? t / G : 0, gr = active ? t - (t / G) * G : 0;
  const int mx0 = kRectT * gc - BB, my0 = kRectT * gr - BB;
  const double dk = 2.0 * M_PI / L;
  int xi[kRectT], yi[kRectT];
#pragma unroll
  for (int j = 0; j < kRectT; ++j) {
    xi[j] = min(max(mx0 + j + B, 0), nk - 1);
    yi[j] = min(max(my0 + j + B, 0), nk - 1);
  }
  double ar[kRectT][kRectT], ai[kRectT][kRectT];
#pragma unroll
  for (int i = 0; i < kRectT; ++i)
#pragma unroll
    for (int j = 0; j < kRectT; ++j) ar[i][j] = ai[i][j] = 0.0;
  for (int64_t cb = e0; cb < e1; cb += chunk) {
    const int ne = (int)min((int64_t)chunk, e1 - cb);
    __syncthreads();  // the previous chunk's tables are consumed
    for (int q = threadIdx.x; q < 2 * ne * nk; q += blockDim.x) {
      const int axis = q / (ne * nk), r = q - axis * ne * nk;
      const int el = r / nk, mi = r - el * nk;
      const double* p = edges + 6 * (cb + el);
      const double k = dk * (mi - B);
      double sv, cv, sh, ch;
      sincos(k * (axis == 0 ? p[2] : p[3]), &sv, &cv);
      sincos(0.5 * k * (axis == 0 ? p[0] : p[1]), &sh, &ch);
      // e^{-i k c}, times the edge weight on the x axis
      double pr = cv, pi = -sv;
      if (axis == 0) {
        const double wr = p[4], wi = p[5];
        const double tr = wr * pr - wi * pi;
        pi = wr * pi + wi * pr;
        pr = tr;
      }
      tab[(axis * chunk + el) * nk + mi] = make_double4(pr, pi, ch, sh);
    }
    __syncthreads();
    if (active) {
      for (int el = 0; el < ne; ++el) {
        const double dx = edges[6 * (cb + el)], dy = edges[6 * (cb + el) + 1];
        double4 yv[kRectT];
#pragma unroll
        for (int i = 0; i < kRectT; ++i) yv[i] = tab[(chunk + el) * nk + yi[i]];
#pragma unroll
        for (int j = 0; j < kRectT; ++j) {
          const double4 xv = tab[el * nk + xi[j]];  // one column at a time (register budget)
          const double kx = dk * (mx0 + j);
#pragma unroll
          for (int i = 0; i < kRectT; ++i) {
            const double ky = dk * (my0 + i);
            const double kn = kx * dy - ky * dx;
            const double u = 0.5 * (kx * dx + ky * dy);
            const double su = xv.w * yv[i].z + xv.z * yv[i].w;
            double r = __builtin_amdgcn_rcp(u);
            r = fma(r, fma(-u, r, 1.0), r);
            r = fma(r, fma(-u, r, 1.0), r);
            const double sc = fabs(u) < 1e-4 ? 1.0 - u * u * (1.0 / 6.0) : su * r;
            const double a = kn * sc;
            const double hr = xv.x * yv[i].x - xv.y * yv[i].y;
            const double hi = xv.x * yv[i].y + xv.y * yv[i].x;
            ar[i][j] = fma(a, hr, ar[i][j]);
            ai[i][j] = fma(a, hi, ai[i][j]);
          }
        }
      }
    }
  }
  if (!active) return;
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
  if (B < 0 || BB < B || B > kPolyMaxBand || BB > 1024) return hipErrorInvalidValue;
  const int W = 2 * BB + 1, nk = 2 * B + 1;
  const int G = (W + kRectT - 1) / kRectT;
  const int wpc = (G * G + 511) / 512;
  const int threads = ((((G * G + wpc - 1) / wpc) + 63) / 64) * 64;
  if ((int64_t)wpc * nclips >= (1ll << 31)) return hipErrorInvalidValue;
  // edges per chunk: two tables of nk 32-byte entries per edge within 48 KiB of LDS
  const int chunk = std::max(1, (48 * 1024) / (2 * nk * (int)sizeof(double4)));
  const size_t lds = (size_t)2 * chunk * nk * sizeof(double4);
  hipLaunchKernelGGL(k_poly_band, dim3((unsigned)(wpc * nclips)), dim3(threads), lds, s, edges,
                     eoff, pdc, n, L, B, BB, chunk, wpc, Mq, Mp);
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
