// Launchers of the device kernel-set construction (kbuild_kernels.hip).  All
// pointers are device pointers; every launcher is asynchronous on `s`.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace huvk {

// Pupil model of one optics: projection optics (euv = 0, amplitude 1, medium
// index n_medium) or EUV mirror optics (euv = 1, central obscuration as a
// fraction of NA, amplitude sqrt(transmission), vacuum).
struct PupilParams {
  int euv;
  int paraxial;
  double na;
  double n_medium;
  double obscuration;
  double amplitude;
};

// A[s][r] = sqrt(w_s) P(cfx[r] + sx_s, cfy[r] + sy_s; z, lambda), pts = (sx, sy, w) triples;
// A is n_pts columns of n_cand values.
hipError_t launch_pupil_factor(const double* cfx, const double* cfy, int64_t n_cand,
                               const double* pts, int n_pts, const PupilParams& pp, double z,
                               double lambda, double2* A, hipStream_t s);
// Per column: sum |A|^2, |A[s][0]|^2, max |A|^2 / w_s, nonzero flag; per row
// (row_nz may be null): nonzero flag.
hipError_t launch_factor_stats(const double2* A, int64_t n_cand, const double* pts, int n_pts,
                               double* col_sum, double* col_dc, double* col_peak, int* col_nz,
                               int* row_nz, hipStream_t s);
// Split-K count used by launch_gram for a D x D output summed over L terms.
int gram_splits(int D, int64_t L);
// h = P^H P (D x D column-major, both triangles, real diagonal) of the column
// vectors of A (mode 0: D = n_pts, L = n_cand) or of its conjugated rows
// (mode 1: D = n_cand, L = n_pts); part: splits x D x D scratch.
hipError_t launch_gram(const double2* A, int64_t n_cand, int mode, int D, int64_t L, int splits,
                       double2* part, double2* h, hipStream_t s);
// U[c] = A V[c] inv_sqrt[c] over the kept columns keep[0..ne), c < K.
hipError_t launch_kernels_from_vectors(const double2* A, int64_t n_cand, const double2* V,
                                       const double* inv_sqrt, const int* keep, int ne, int K,
                                       double2* U, hipStream_t s);

}  // namespace huvk
