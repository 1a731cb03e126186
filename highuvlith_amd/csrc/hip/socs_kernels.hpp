// Launchers of the SOCS image kernels (socs_kernels.hip).  All pointers are
// device pointers; every launcher is asynchronous on `stream`.
#pragma once

#include <hip/hip_runtime.h>

namespace huvk {

// Coarse grid for band half-width B on an n-grid: the smallest supported m
// (32, 64, 128, 192, 256) with m >= 4B + 1 and m <= n; 0 if none.
int choose_coarse(int n, int B);
// x classes (and y residues) of the 32-point field transforms at coarse size m.
int coarse_classes(int m);
// Band half-width the field kernel is compiled for at coarse size m: (m - 1) / 4.
// Kernel boxes and band spectra are stored at this width.
int coarse_band(int m);

// Band FFTs of length n run as four-step N1 x (n / N1): band_fft_default(n)
// is the built-in N1, band_fft_options lists the supported ones (returns the
// count).  n1 = 0 selects the default.
int band_fft_default(int n);
int band_fft_options(int n, int* n1s);
// x pass of the field kernel: xlanes = 1 (one lane per band row, every
// column) or 2 (two lanes per row, half of the bins each; m >= 64, reads the
// column-paired layouts); xnb = 4, 6 or 8 columns per load block.  Zero
// selects fields_lanes_default(m) and kFieldsXnbDefault / kFieldsXnbDefault2.
constexpr int kFieldsXnbDefault = 6;
constexpr int kFieldsXnbDefault2 = 6;
int fields_lanes_default(int m);
bool fields_lanes_supported(int m, int xlanes);

hipError_t launch_rows_band(int n, int n1, int B, int nclips, const float* masks,
                            const float2* tw_n, float2* tmp, hipStream_t s);
// Mq: [nclips][W][W] band spectrum (W = 2BB+1, [kx + BB][ky + BB]), zero
// beyond |k| <= B.  Mp (optional, may be null): the same values column-paired,
// [nclips][(W + 1) / 2][W][2] with column kx + BB at [(kx + BB + 1) / 2][.][(kx + BB + 1) % 2]
// and a zero column in front (the layout of the two-lane field kernel).
hipError_t launch_cols_band(int n, int n1, int B, int BB, int nclips, const float2* tmp,
                            const float2* tw_n, const float* sinc_t, float2* Mq, float2* Mp,
                            hipStream_t s);
// Complex rasters (masks: [nclips][n][n] interleaved complex): every row is
// transformed (tmp: [nclips][2B+1][n]) and both signs of kx are column
// passes; Mq / Mp as for launch_cols_band.
hipError_t launch_rows_band_c(int n, int n1, int B, int nclips, const float2* masks,
                              const float2* tw_n, float2* tmp, hipStream_t s);
hipError_t launch_cols_band_c(int n, int n1, int B, int BB, int nclips, const float2* tmp,
                              const float2* tw_n, const float* sinc_t, float2* Mq, float2* Mp,
                              hipStream_t s);
// kern: [Q][nplanes][Kmax] boxes (W = 2 coarse_band(m) + 1, Q = m / 32; copy q
// multiplied by e^{2 pi i kx q / m}), lam: [nplanes][Kmax]; Mq: one box per
// clip; coarse: [nclips][nplanes][m][Q][32], row y, column x = Q j + q at
// [y][q][j] (class-major).  Boxes are W x W ([kx + BB][ky + BB])
// for xlanes = 1 and column-paired (W + 1) x W (launch_cols_band's Mp) for 2.
// S kernel splits (fields_splits): coarse then holds S partial image sets,
// [S][nclips][nplanes][m][m], summed by launch_coarse_rows.
int fields_splits(int m, int nclips, int Kmax);
hipError_t launch_socs_fields(int m, int xnb, int xlanes, int nclips, int nplanes, int Kmax, int S,
                              const float2* kern, const float* lam, const float2* Mq,
                              const float2* tw_m, float* coarse, hipStream_t s);
hipError_t launch_coarse_rows(int m, int B, int nimg, int S, const float* coarse,
                              const float2* tw_m, float2* tmp1, hipStream_t s);
hipError_t launch_coarse_cols(int m, int B, int nimg, float flare, const float2* tmp1,
                              const float2* tw_m, float2* Ib, hipStream_t s);
// coarse_rows + coarse_cols of each image in one work-group (row spectra in
// LDS), for m = 64, 128, 192; the same Ib as the two launches.
bool coarse_fused_supported(int m);
hipError_t launch_coarse_fused(int m, int B, int nimg, int S, const float* coarse, float flare,
                               const float2* tw_m, float2* Ib, hipStream_t s);
// Ib: [nimg][2G+1][G+1] Hermitian half band (gy + G, gx).
hipError_t launch_band_image(int n, int G, int nimg, const float2* Ib, const float2* tw_n,
                             float* out, hipStream_t s);

// The ILT fidelity term evaluated where band_image would store the image
// (images img = target * ncond + condition): r = sigma(steepness (dose I -
// threshold)) - Z (cost_kind 1) or dose I - Z (0), Z = target row
// trows[target] (trows null: target); partial[img][p] = sum of r^2 over the
// image's row class p (n / 32 per image, FP64, fixed order); G (optional) =
// 2 r dr/dI inv_n per pixel.
struct FidelityEpi {
  const float* target = nullptr;
  const int* trows = nullptr;
  int ncond = 1;
  int cost_kind = 0;
  float threshold = 0.f, steepness = 0.f, inv_n = 0.f;
  const float* dose = nullptr;
  float* G = nullptr;
  double* partial = nullptr;
};
hipError_t launch_band_image_fidelity(int n, int G, int nimg, const float2* Ib,
                                      const float2* tw_n, const FidelityEpi& epi, hipStream_t s);

// The ILT chain rule evaluated where band_image would store the image: image
// j is dC/dm of ILT row j, and row (orows ? orows[j] : j) of grad receives
// dC/dtheta = (dC/dm + reg) beta m (1 - m) (m, reg: rows j; reg optional).
struct ChainEpi {
  const float* m = nullptr;
  const float* reg = nullptr;
  const int* orows = nullptr;
  float beta = 0.f;
  float* grad = nullptr;
};
hipError_t launch_band_image_chain(int n, int G, int nimg, const float2* Ib, const float2* tw_n,
                                   const ChainEpi& ch, hipStream_t s);

}  // namespace huvk
