// Launchers of the ILT kernels (ilt_kernels.hip); device pointers, async on `s`.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace huvk {

hipError_t launch_ilt_mask(const float* theta, float* m, float* t, int64_t total, float beta,
                           float a, hipStream_t s);
hipError_t launch_ilt_pointwise(const float* image, const float* target, int nimg, int ncond,
                                int64_t npix, int bpi, int cost_kind, float threshold,
                                float steepness, const float* dose, float* G, double* partial,
                                hipStream_t s);
hipError_t launch_ilt_reg(const float* m, int n, int ntargets, int bpi, float tv_w, float tv_eps,
                          float bin_w, float* reg_grad, double* partial, hipStream_t s);
hipError_t launch_ilt_cost_reduce(const double* fid, const double* reg, int ntargets, int ncond,
                                  int bpi, int reg_bpi, const float* weight, double inv_n,
                                  double* cost, hipStream_t s);
hipError_t launch_ilt_adjoint(int m, int nimg, const float2* Mq, const float2* Gb, int B,
                              int ncond, const int* cond_K, const float2* const* cond_kern,
                              const float* const* cond_lam, float lam_scale, float g0_scale,
                              float flare, float2* Rout, hipStream_t s);
hipError_t launch_ilt_hband(const float2* Rin, int B, int ntargets, int ncond, const float* weight,
                            float scale, float2* H, hipStream_t s);
hipError_t launch_ilt_chain(const float* dcdm, const float* reg, const float* m, float beta,
                            int64_t total, float* grad, hipStream_t s);
// mode 0: per-target sum x.y ; mode 1: per-target max |x|.  part: ntargets x nblk.
hipError_t launch_vec_reduce(const float* x, const float* y, int64_t npix, int ntargets, int nblk,
                             int mode, double* part, double* out, hipStream_t s);
hipError_t launch_vec_axpy(const float* x, const float* d, const double* coef, int64_t npix,
                           int ntargets, float* y, hipStream_t s);
hipError_t launch_vec_dir(const float* g, const double* a, const double* b, int64_t npix,
                          int ntargets, float* d, hipStream_t s);
// dst[j] = src[idx[j]] (gather) / dst[idx[j]] = src[j] (scatter), rows of npix.
hipError_t launch_rows_gather(const float* src, const int* idx, int64_t npix, int nrows,
                              float* dst, hipStream_t s);
hipError_t launch_rows_scatter(const float* src, const int* idx, int64_t npix, int nrows,
                               float* dst, hipStream_t s);
hipError_t launch_vec_select(const float* src, const int* sel, int64_t npix, int ntargets,
                             float* dst, hipStream_t s);

}  // namespace huvk
