// Host-side (FP64) half of the imaging engine: optics, pupil fills, adaptive
// source sampling, TCC factor assembly and its Hermitian eigendecomposition,
// SOCS kernel selection, and mask rasterisation.  This is the part the
// reference keeps on the CPU once per optical setting
// (crates/highuvlith-core/src/aerial.rs:941-1317, math/linalg.rs:181-679);
// the per-image work runs on the GPU behind include/huv.h.
#pragma once

#include <complex>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace huv {

using cd = std::complex<double>;

// Error codes mirror LithographyError (crates/highuvlith-core/src/error.rs:3-61).
enum Status : int {
  OK = 0,
  INVALID_PARAMETER = 1,
  DIMENSION_MISMATCH = 2,
  NUMERICAL_ERROR = 3,
  INTERNAL_ERROR = 4,
  DEVICE_ERROR = 5,
  NO_DIFFRACTION_ORDERS = 6,
  PUPIL_SAMPLING_TOO_DENSE = 7,
};

// ---------------------------------------------------------------------------
// Optics (optics/mod.rs:201-360, optics/euv.rs:49-165)
// ---------------------------------------------------------------------------
enum OpticsKind : int { OPTICS_PROJECTION = 0, OPTICS_EUV = 1 };

struct Optics {
  int kind = OPTICS_PROJECTION;
  double na = 0.75;
  double immersion_index = 1.0;     // projection optics only
  double flare = 0.02;              // flare fraction phi
  double central_obscuration = 0.0; // EUV only (fraction of NA)
  double transmission = 1.0;        // EUV only (intensity)
  int paraxial_defocus = 0;
};

// Defocus phase of the order at NA*rho in a medium of index n
// (optics/mod.rs:155-172).
double defocus_phase_in_medium(double z_nm, double na_rho, double n_medium,
                               double lambda_nm, bool paraxial);
// Pupil P(fx, fy; z, lambda) at frequencies normalised to NA/lambda.
cd pupil_function(const Optics& o, double fx, double fy, double z_nm,
                  double lambda_nm);

// ---------------------------------------------------------------------------
// Pupil fills (source.rs:163-186, 1178-1248) plus the annular-sector fills
// the lithography workloads use (x-dipole, quasar).
// ---------------------------------------------------------------------------
enum IllumKind : int {
  ILLUM_CONVENTIONAL = 0,     // p0 = sigma
  ILLUM_ANNULAR = 1,          // p0 = sigma_inner, p1 = sigma_outer
  ILLUM_QUADRUPOLE = 2,       // p0 = sigma_center, p1 = sigma_radius, p2 = opening_deg
  ILLUM_DIPOLE = 3,           // p0 = sigma_center, p1 = sigma_radius, p2 = orientation_deg
  ILLUM_GAUSSIAN = 4,         // p0 = sigma
  ILLUM_SECTOR_POLES = 5,     // p0 = sigma_inner, p1 = sigma_outer, p2 = opening_deg,
                              // p3 = first pole angle (deg), p4 = number of poles
};

struct Illumination {
  int kind = ILLUM_CONVENTIONAL;
  double p[6] = {0.5, 0, 0, 0, 0, 0};
};

double illumination_at(const Illumination& s, double sx, double sy);

// ---------------------------------------------------------------------------
// Source sampling (aerial.rs:372-947)
// ---------------------------------------------------------------------------
struct SourcePoint {
  double sx, sy, weight;
};

using Fill = std::function<double(double, double)>;

// points_per_axis <= 0: adaptive rotated-lattice sampling (the default).
int sample_intensity(const Fill& f, int points_per_axis, double df_sigma,
                     std::vector<SourcePoint>& out, std::string& err);

// ---------------------------------------------------------------------------
// Dense complex linear algebra (math/linalg.rs)
// ---------------------------------------------------------------------------
// Column-major complex matrix.
struct ColMatrix {
  int64_t rows = 0, cols = 0;
  std::vector<cd> data;
  cd* col(int64_t c) { return data.data() + c * rows; }
  const cd* col(int64_t c) const { return data.data() + c * rows; }
  double frobenius_norm_sq() const;
  // Hermitian Gram matrix X^H X (cols x cols), column-major, both triangles.
  std::vector<cd> gram() const;
};

// Complete eigendecomposition of a Hermitian matrix (n x n column-major):
// eigenvalues decreasing, eigenvectors as columns (n x count).  Householder
// tridiagonalisation + phase normalisation + implicit QL (linalg.rs:324-551).
bool hermitian_eigen_desc(const std::vector<cd>& h, int64_t n, int64_t count,
                          std::vector<double>& values, std::vector<cd>& vectors);

// Leading eigenpairs of X X^H by randomized subspace iteration
// (linalg.rs:657-679).  Gaussian stream: splitmix64 + Box-Muller (the
// reference draws from rand 0.9 StdRng + rand_distr; that stream is not
// reproduced, only the algorithm).
void randomized_eigen(const ColMatrix& x, int64_t l, int power_iters, uint64_t seed,
                      std::vector<double>& values, ColMatrix& u);

// ---------------------------------------------------------------------------
// SOCS kernel sets (aerial.rs:941-1317)
// ---------------------------------------------------------------------------
enum Method : int { METHOD_AUTO = -1, METHOD_DENSE_TCC = 0, METHOD_DENSE_GRAM = 1,
                    METHOD_RANDOMIZED = 2 };

int select_kernel_count(const double* values, int64_t nvals, double trace,
                        int64_t max_kernels, double fraction);

struct KernelSet {
  int n = 0;                        // grid size
  double pixel_nm = 0;
  std::vector<int64_t> support;     // flat row-major indices i*n + j
  std::vector<int64_t> nonzero_rows;
  std::vector<double> eigenvalues;  // absolute, decreasing
  std::vector<cd> values;           // K x S, kernel-major
  double captured_energy_fraction = 0;
  double clear_field = 0;
  double pupil_peak = 0;
  double defocus_nm = 0, wavelength_nm = 0;
  int method = 0;
  int64_t num_columns = 0;
  bool support_exceeds_nyquist = false;
  int num_kernels() const { return (int)eigenvalues.size(); }
  int64_t num_support() const { return (int64_t)support.size(); }
  // ClearField normalisation divisor, or 1 for dark-field sets (aerial.rs:1727-1733).
  double normalization_divisor() const;
};

struct BuildSettings {
  int max_kernels = 48;
  double kernel_energy_fraction = 1.0;
  int method = METHOD_AUTO;
};

int build_kernel_set(const Optics& optics, const std::vector<SourcePoint>& points,
                     int n, double pixel_nm, double defocus_nm, double wavelength_nm,
                     const BuildSettings& settings, KernelSet& out, std::string& err);

// The stages of build_kernel_set that the device build shares with it.
// Candidate mask frequencies |f| <= (1 + sigma_max) NA/lambda in row-major
// grid order (candidate 0 is DC), in units of the cutoff NA/lambda
// (aerial.rs:1053-1075).
struct Candidates {
  std::vector<int64_t> idx;     // flat grid index i * n + j
  std::vector<double> fx, fy;   // frequency / cutoff
  bool exceeds_nyquist = false;
};
int kernel_candidates(const Optics& optics, const std::vector<SourcePoint>& points, int n,
                      double pixel_nm, double wavelength_nm, Candidates& out, std::string& err);
// Decomposition path for an n_cand x n_cols factor (aerial.rs:1118-1128).
int choose_build_method(int64_t n_cand, int64_t n_cols, const BuildSettings& settings);
// Fill `out` from the chosen kernels (values K x support, kernel-major).
int assemble_kernel_set(int n, double pixel_nm, double defocus_nm, double wavelength_nm,
                        std::vector<int64_t> support, std::vector<double> eig,
                        std::vector<cd> values, double trace, double clear_field,
                        double pupil_peak, int method, int64_t n_eff, bool exceeds_nyquist,
                        KernelSet& out, std::string& err);

// ---------------------------------------------------------------------------
// Mask rasterisation: axis-aligned rectangles painted in order over a
// background, exact area coverage per pixel (mask.rs:590, 1365-1571).
// ---------------------------------------------------------------------------
struct RectFeature {
  double x, y, w, h;  // centre and size (nm)
  cd value;           // complex amplitude inside
};

void rasterize_rects(const std::vector<RectFeature>& rects, cd background, int n,
                     double pixel_nm, std::vector<cd>& out);

// A disjoint rectangle [x0, x1) x [y0, y1) (nm) of weight w = value - background.
struct RectCell {
  double x0, x1, y0, y1;
  cd w;
};
// Painter's-rule resolution of rectangles into disjoint cells (see the
// definition); false when the edge grid would exceed max_cells cells.
bool resolve_rect_cells(const std::vector<RectFeature>& rects, cd background, int n,
                        double pixel_nm, int64_t max_cells, std::vector<RectCell>& out);

}  // namespace huv
