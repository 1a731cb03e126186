// Launchers of the SOCS image kernels (socs_kernels.hip).  All pointers are
// device pointers; every launcher is asynchronous on `stream`.
#pragma once

#include <hip/hip_runtime.h>

namespace huvk {

// Coarse grid for band half-width B on an n-grid: the smallest supported m
// (32, 64, 128, 192, 256) with m >= 4B + 1 and m <= n; 0 if none.
int choose_coarse(int n, int B);
// x classes (and y residues) used for coarse size m.
int coarse_classes(int m);

hipError_t launch_rows_band(int n, int B, int nclips, const float* masks, const float2* tw_n,
                            float2* tmp, hipStream_t s);
hipError_t launch_cols_band(int n, int B, int m, int nclips, const float2* tmp,
                            const float2* tw_n, const float2* tw_m, const float* sinc_t,
                            float2* Mq, hipStream_t s);
hipError_t launch_socs_fields(int m, int B, int nclips, int nplanes, const int* plane_K,
                              const float2* const* plane_kern, const float* const* plane_lam,
                              const float2* Mq, float* coarse, hipStream_t s);
hipError_t launch_coarse_rows(int m, int B, int nimg, const float* coarse, const float2* tw_m,
                              float2* tmp1, hipStream_t s);
hipError_t launch_coarse_cols(int m, int B, int nimg, float flare, const float2* tmp1,
                              const float2* tw_m, float2* Ib, hipStream_t s);
hipError_t launch_band_image(int n, int G, int nimg, const float2* Ib, const float2* tw_n,
                             float* out, hipStream_t s);

}  // namespace huvk
