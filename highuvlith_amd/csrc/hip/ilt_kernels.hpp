// Launchers of the ILT kernels (ilt_kernels.hip); device pointers, async on `s`.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace huvk {

// m and t (real absorber ar, or tc for a complex one ar + i ai; each may be
// null) of count evaluation rows read in place: row j's theta is
// theta[rows[j]] + coef[j] dir[rows[j]] (rows null: j; coef null: theta only).
hipError_t launch_ilt_mask_rows(const float* theta, const int* rows, const float* dir,
                                const double* coef, int64_t npix, int count, float beta, float ar,
                                float ai, float* m, float* t, float2* tc, hipStream_t s);
hipError_t launch_ilt_reg(const float* m, int n, int ntargets, int bpi, float tv_w, float tv_eps,
                          float bin_w, float* reg_grad, double* partial, hipStream_t s);
hipError_t launch_ilt_cost_reduce(const double* fid, const double* reg, int ntargets, int ncond,
                                  int bpi, int reg_bpi, const float* weight, double inv_n,
                                  double* cost, hipStream_t s);
// Adjoint band of nimg = ntargets x ncond images on coarse grid m (32 or 64):
// G0g scratch [nimg][m][m] floats, Rpart [nimg][KS][2B+1][2B+1] partial bands
// (the kernels of each condition dealt over KS work-groups per image).
hipError_t launch_ilt_adjoint(int m, int nimg, int KS, const float2* Mq, const float2* Gb, int B,
                              int ncond, const int* cond_K, const float2* const* cond_kern,
                              const float* const* cond_lam, float lam_scale, float g0_scale,
                              float flare, float* G0g, float2* Rpart, hipStream_t s);
// The same adjoint for coarse grids m = 128, 192, 256: team FFTs through HBM
// (g0tmp: nimg x (4B+1) x m complex; buf: nimg x KS x (2B+1) x m complex;
// tw_m: the W_m table); Rpart holds one band per kernel (KS = the largest
// kernel count, zero bands past a condition's own count).
hipError_t launch_ilt_adjoint_big(int m, int nimg, int KS, const float2* Mq, const float2* Gb,
                                  int B, int ncond, const int* cond_K,
                                  const float2* const* cond_kern, const float* const* cond_lam,
                                  float lam_scale, float g0_scale, float flare,
                                  const float2* tw_m, float2* g0tmp, float* G0g, float2* buf,
                                  float2* Rpart, hipStream_t s);
// H = sum_c w_c (conj(c) R_c(f) + c conj(R_c(-f))) / n^2, c = cr + i ci = 1 - absorber.
hipError_t launch_ilt_hband(const float2* Rin, int B, int ntargets, int ncond, int KS,
                            const float* weight, float cr, float ci, float inv_n2, float2* H,
                            hipStream_t s);
// Density-filter band for band_image: Ib[t][gy + Bf][gx] = X[gx + Bf][gy + Bf]
// exp(-c (gx^2 + gy^2)) inv_n2 for count boxes X of half-width Bf.
hipError_t launch_ilt_filter_band(const float2* X, int Bf, float c, float inv_n2, int count,
                                  float2* Ib, hipStream_t s);
// Local-proxy dC/dm from G (ntargets x ncond images); part: ntargets x ncond x bpi.
hipError_t launch_ilt_proxy(const float* G, int ntargets, int ncond, int64_t npix, int bpi,
                            const float* weight, float flare, double* part, float* dcdm,
                            hipStream_t s);
// dC/dtheta of count evaluation rows into grad rows orows[j] (null: j).
hipError_t launch_ilt_chain(const float* dcdm, const float* reg, const float* m, float beta,
                            int64_t npix, int count, const int* orows, float* grad,
                            hipStream_t s);
// Per-target vector operations on `count` target rows listed in `rows`
// (device; null = rows 0..count-1); other rows are left untouched.
// Up to nops <= 3 per-target reductions of a and b in one pass, op r in bits
// 4r..4r+3 of ops: 0 a.a, 1 a.b, 2 b.b, 3 max |a| (b may be null for 0 and
// 3).  part: nops x T x nblk; out[r * T + t] for the listed t (T: the stride,
// at least count); out null: the partials only (summed by the caller, over
// the blocks in order).
hipError_t launch_vec_reduce(const float* a, const float* b, int64_t npix, int count,
                             const int* rows, int nblk, int ops, int nops, int T, double* part,
                             double* out, hipStream_t s);
// d = a[t] g + b[t] d
hipError_t launch_vec_dir(const float* g, const double* a, const double* b, int64_t npix,
                          int count, const int* rows, float* d, hipStream_t s);
// launch_vec_dir, and in the same pass the reductions max |d| (r = 0) and d.g
// (r = 1) of the new d over the listed rows: part as launch_vec_reduce with
// ops (3, 1) would write it, bit for bit.
hipError_t launch_vec_dir_reduce(const float* g, const double* a, const double* b, int64_t npix,
                                 int count, const int* rows, int nblk, int T, float* d,
                                 double* part, hipStream_t s);
// dst[j] = src[idx[j]] (gather) / dst[idx[j]] = src[j] (scatter), rows of npix.
hipError_t launch_rows_gather(const float* src, const int* idx, int64_t npix, int nrows,
                              float* dst, hipStream_t s);
hipError_t launch_rows_scatter(const float* src, const int* idx, int64_t npix, int nrows,
                               float* dst, hipStream_t s);
// out[j] = x[rows[j]] + coef[j] d[rows[j]], j < count (compact line-search trials)
hipError_t launch_vec_axpy_gather(const float* x, const float* d, const double* coef, int64_t npix,
                                  int count, const int* rows, float* out, hipStream_t s);
// x[rows[j]] += coef[j] d[rows[j]] in place (accepted line-search steps)
hipError_t launch_vec_axpy_rows(float* x, const float* d, const double* coef, int64_t npix,
                                int count, const int* rows, hipStream_t s);

}  // namespace huvk
