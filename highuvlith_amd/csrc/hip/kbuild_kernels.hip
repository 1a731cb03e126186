// Device half of the SOCS kernel-set construction (FP64, gfx950)
// (crates/highuvlith-core/src/aerial.rs:1041-1317, math/linalg.rs:181-232):
//   factor  A[s][r] = sqrt(w_s) P(f_r / f_c + s_s; z, lambda)   (one column per source point)
//   stats   per column: sum |A|^2, |A[s][DC]|^2, max |A|^2 / w_s, any nonzero
//   Gram    G = P^H P of the column vectors (mode 0, the dense-Gram path) or
//           of the conjugated rows (mode 1, the dense-TCC path), FP64 with
//           split-K partials summed in a fixed order
//   kernels U[c] = A V[c] / sqrt(lambda_c) from the host eigenvectors
// The Hermitian eigendecomposition stays on the host.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kbuild_kernels.hpp"

namespace huvk {

namespace {
__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 zconjmul(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
}  // namespace

// Pupil P(fx, fy; z, lambda) at frequencies normalised to NA/lambda
// (optics/mod.rs:155-172, 310-355; optics/euv.rs:129-165).
__device__ double2 pupil_dev(const PupilParams& p, double fx, double fy, double z, double lambda) {
  const double rho2 = fx * fx + fy * fy;
  if (rho2 > 1.0) return make_double2(0.0, 0.0);
  const double rho = sqrt(rho2);
  if (p.euv && rho < p.obscuration) return make_double2(0.0, 0.0);
  double phase = 0.0;
  if (z != 0.0) {
    const double s = p.na * rho / p.n_medium, s2 = s * s;
    if (p.paraxial) {
      phase = M_PI * z * p.n_medium * s2 / lambda;
    } else {
      const double cos_t = sqrt(fmax(1.0 - s2, 0.0));
      phase = 2.0 * M_PI * p.n_medium * z * s2 / ((1.0 + cos_t) * lambda);
    }
  }
  double sn, cs;
  sincos(phase, &sn, &cs);
  return make_double2(p.amplitude * cs, p.amplitude * sn);
}

__global__ void k_pupil_factor(const double* __restrict__ cfx, const double* __restrict__ cfy,
                               int64_t n_cand, const double* __restrict__ pts, int n_pts,
                               PupilParams pp, double z, double lambda,
                               double2* __restrict__ A) {
  const int64_t total = n_cand * n_pts;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = i / n_cand, r = i - s * n_cand;
    const double sx = pts[3 * s], sy = pts[3 * s + 1], w = pts[3 * s + 2];
    const double2 v = pupil_dev(pp, cfx[r] + sx, cfy[r] + sy, z, lambda);
    const double sw = sqrt(w);
    A[i] = make_double2(v.x * sw, v.y * sw);
  }
}

// One work-group per column s: sum |A|^2 (tree order), |A[s][0]|^2, max |A|^2 / w_s, nonzero.
__global__ __launch_bounds__(256) void k_factor_col_stats(const double2* __restrict__ A,
                                                          int64_t n_cand,
                                                          const double* __restrict__ pts,
                                                          double* __restrict__ col_sum,
                                                          double* __restrict__ col_dc,
                                                          double* __restrict__ col_peak,
                                                          int* __restrict__ col_nz) {
  __shared__ double ssum[256], speak[256];
  __shared__ int snz[256];
  const int s = blockIdx.x;
  const double2* a = A + (size_t)s * n_cand;
  const double w = pts[3 * s + 2];
  double sum = 0.0, peak = 0.0;
  int nz = 0;
  for (int64_t r = threadIdx.x; r < n_cand; r += 256) {
    const double2 v = a[r];
    const double q = v.x * v.x + v.y * v.y;
    sum += q;
    peak = fmax(peak, q / w);
    nz |= (v.x != 0.0 || v.y != 0.0) ? 1 : 0;
  }
  ssum[threadIdx.x] = sum;
  speak[threadIdx.x] = peak;
  snz[threadIdx.x] = nz;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      ssum[threadIdx.x] += ssum[threadIdx.x + h];
      speak[threadIdx.x] = fmax(speak[threadIdx.x], speak[threadIdx.x + h]);
      snz[threadIdx.x] |= snz[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    col_sum[s] = ssum[0];
    col_peak[s] = speak[0];
    col_nz[s] = snz[0];
    const double2 d = a[0];
    col_dc[s] = d.x * d.x + d.y * d.y;
  }
}

// row_nz[r] = any nonzero A[s][r] over the columns.
__global__ void k_factor_row_nz(const double2* __restrict__ A, int64_t n_cand, int n_pts,
                                int* __restrict__ row_nz) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n_cand;
       r += (int64_t)gridDim.x * blockDim.x) {
    int nz = 0;
    for (int s = 0; s < n_pts && !nz; ++s) {
      const double2 v = A[(size_t)s * n_cand + r];
      nz = (v.x != 0.0 || v.y != 0.0) ? 1 : 0;
    }
    row_nz[r] = nz;
  }
}

// Gram tiles.  Vector a, element k: mode 0 A[a][k] (column a of the factor,
// k < n_cand), mode 1 conj(A[k][a]) (row a, k < n_pts).  Work-group (tile
// pair p, split j): the 16 x 16 outputs a in tile ta, b in tile tb (ta <= tb)
// summed over k in [j kper, (j + 1) kper), staged 32 k at a time in LDS.
constexpr int kGT = 16, kGK = 32;
__global__ __launch_bounds__(256) void k_gram_tiles(const double2* __restrict__ A, int64_t n_cand,
                                                    int mode, int D, int64_t L, int64_t kper,
                                                    double2* __restrict__ part) {
  __shared__ double2 va[kGK][kGT + 1], vb[kGK][kGT + 1];
  // decode the upper-triangular tile pair
  int p = blockIdx.x, ta = 0;
  const int nt = (D + kGT - 1) / kGT;
  while (p >= nt - ta) {
    p -= nt - ta;
    ++ta;
  }
  const int tb = ta + p;
  const int split = blockIdx.y;
  const int64_t k0 = (int64_t)split * kper, k1 = min(L, k0 + kper);
  const int tx = threadIdx.x % kGT, ty = threadIdx.x / kGT;
  auto elem = [&](int a, int64_t k) -> double2 {
    if (mode == 0) return A[(size_t)a * n_cand + k];
    const double2 v = A[(size_t)k * n_cand + a];
    return make_double2(v.x, -v.y);
  };
  double2 acc = make_double2(0.0, 0.0);
  for (int64_t kc = k0; kc < k1; kc += kGK) {
    for (int e = threadIdx.x; e < kGK * kGT; e += 256) {
      // mode 0 walks k fastest (contiguous columns), mode 1 walks a fastest (contiguous rows)
      const int kk = mode == 0 ? e % kGK : e / kGT;
      const int i = mode == 0 ? e / kGK : e % kGT;
      const int64_t k = kc + kk;
      const int a = ta * kGT + i, b = tb * kGT + i;
      va[kk][i] = (k < k1 && a < D) ? elem(a, k) : make_double2(0.0, 0.0);
      vb[kk][i] = (k < k1 && b < D) ? elem(b, k) : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll 8
    for (int kk = 0; kk < kGK; ++kk) {
      const double2 u = zconjmul(va[kk][ty], vb[kk][tx]);
      acc.x += u.x;
      acc.y += u.y;
    }
    __syncthreads();
  }
  const int a = ta * kGT + ty, b = tb * kGT + tx;
  if (a < D && b < D) part[((size_t)split * D + a) * D + b] = acc;
}

// h (column-major D x D, both triangles): h[(a, b)] = sum over splits in order
// for a <= b, h[(b, a)] = conj, real diagonal (linalg.rs:181-232 layout).
__global__ void k_gram_sum(const double2* __restrict__ part, int D, int splits,
                           double2* __restrict__ h) {
  const int64_t total = (int64_t)D * D;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int a = (int)(i / D), b = (int)(i - (int64_t)a * D);
    if (a > b) continue;
    double2 v = make_double2(0.0, 0.0);
    for (int j = 0; j < splits; ++j) {
      const double2 w = part[((size_t)j * D + a) * D + b];
      v.x += w.x;
      v.y += w.y;
    }
    if (a == b) v.y = 0.0;
    h[(size_t)b * D + a] = v;
    h[(size_t)a * D + b] = make_double2(v.x, -v.y);
  }
}

// U[c][r] = sum_s (V[c][s] / sqrt(lambda_c)) A[keep[s]][r], s in order.
__global__ void k_kernels_from_vectors(const double2* __restrict__ A, int64_t n_cand,
                                       const double2* __restrict__ V, const double* __restrict__ inv_sqrt,
                                       const int* __restrict__ keep, int ne, int K,
                                       double2* __restrict__ U) {
  const int64_t total = (int64_t)K * n_cand;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i / n_cand);
    const int64_t r = i - (int64_t)c * n_cand;
    const double is = inv_sqrt[c];
    double2 u = make_double2(0.0, 0.0);
    for (int s = 0; s < ne; ++s) {
      const double2 v = V[(size_t)c * ne + s];
      const double2 b = make_double2(v.x * is, v.y * is);
      const double2 t = zmul(b, A[(size_t)keep[s] * n_cand + r]);
      u.x += t.x;
      u.y += t.y;
    }
    U[i] = u;
  }
}

// ===========================================================================
// Launchers
// ===========================================================================
static unsigned grid_of(int64_t total) {
  return (unsigned)std::min<int64_t>((total + 255) / 256, 1 << 16);
}

hipError_t launch_pupil_factor(const double* cfx, const double* cfy, int64_t n_cand,
                               const double* pts, int n_pts, const PupilParams& pp, double z,
                               double lambda, double2* A, hipStream_t s) {
  hipLaunchKernelGGL(k_pupil_factor, dim3(grid_of(n_cand * n_pts)), dim3(256), 0, s, cfx, cfy,
                     n_cand, pts, n_pts, pp, z, lambda, A);
  return hipGetLastError();
}

hipError_t launch_factor_stats(const double2* A, int64_t n_cand, const double* pts, int n_pts,
                               double* col_sum, double* col_dc, double* col_peak, int* col_nz,
                               int* row_nz, hipStream_t s) {
  hipLaunchKernelGGL(k_factor_col_stats, dim3(n_pts), dim3(256), 0, s, A, n_cand, pts, col_sum,
                     col_dc, col_peak, col_nz);
  if (row_nz)
    hipLaunchKernelGGL(k_factor_row_nz, dim3(grid_of(n_cand)), dim3(256), 0, s, A, n_cand, n_pts,
                       row_nz);
  return hipGetLastError();
}

int gram_splits(int D, int64_t L) {
  const int nt = (D + kGT - 1) / kGT;
  const int64_t pairs = (int64_t)nt * (nt + 1) / 2;
  int64_t sp = (2048 + pairs - 1) / pairs;
  sp = std::max<int64_t>(1, std::min<int64_t>(sp, (L + 511) / 512));
  return (int)sp;
}

hipError_t launch_gram(const double2* A, int64_t n_cand, int mode, int D, int64_t L, int splits,
                       double2* part, double2* h, hipStream_t s) {
  if (D < 1 || L < 1 || splits < 1 || splits > 65535) return hipErrorInvalidValue;
  const int nt = (D + kGT - 1) / kGT;
  const int64_t pairs = (int64_t)nt * (nt + 1) / 2;
  if (pairs >= (1ll << 31)) return hipErrorInvalidValue;
  const int64_t kper = (L + splits - 1) / splits;
  hipLaunchKernelGGL(k_gram_tiles, dim3((unsigned)pairs, splits), dim3(256), 0, s, A, n_cand,
                     mode, D, L, kper, part);
  hipLaunchKernelGGL(k_gram_sum, dim3(grid_of((int64_t)D * D)), dim3(256), 0, s, part, D, splits,
                     h);
  return hipGetLastError();
}

hipError_t launch_kernels_from_vectors(const double2* A, int64_t n_cand, const double2* V,
                                       const double* inv_sqrt, const int* keep, int ne, int K,
                                       double2* U, hipStream_t s) {
  hipLaunchKernelGGL(k_kernels_from_vectors, dim3(grid_of((int64_t)K * n_cand)), dim3(256), 0, s,
                     A, n_cand, V, inv_sqrt, keep, ne, K, U);
  return hipGetLastError();
}

}  // namespace huvk
