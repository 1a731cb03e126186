/*
 * huv.h — C ABI of the MI355X backend for the Hopkins/SOCS aerial-image path
 * and the ILT adjoint of highuvlith-core.
 *
 * Two halves:
 *   - host engine (FP64, CPU): source sampling, SOCS kernel-set construction,
 *     mask rasterisation.  Replaces the CPU-side construction the reference
 *     performs once per optical setting.
 *   - device backend (HIP, gfx950): batched band-limited SOCS images and the
 *     ILT cost / adjoint gradient.  Replaces the per-image hot loops.
 *
 * Conventions (crates/highuvlith-core/src/math/fft2d.rs:11-13,
 * aerial.rs:977-983): row-major (ny, nx) grids; FFT index k maps to the
 * signed frequency k (k < n/2) or k - n; complex values are interleaved
 * (re, im) pairs, the layout of num_complex::Complex<f64> / <f32>.
 *
 * Every entry point returns a huv_status; none panics or throws across the
 * ABI.  huv_last_error() returns a thread-local message for the last failure
 * on the calling thread.  A context serialises its own calls; distinct
 * contexts may be used from distinct threads.  Results are deterministic: no
 * floating-point atomics, fixed summation orders.
 */
#ifndef HUV_H
#define HUV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes mirror LithographyError (crates/highuvlith-core/src/error.rs:3-61). */
typedef int huv_status;
#define HUV_OK 0
#define HUV_INVALID_PARAMETER 1
#define HUV_DIMENSION_MISMATCH 2
#define HUV_NUMERICAL_ERROR 3
#define HUV_INTERNAL_ERROR 4
#define HUV_DEVICE_ERROR 5
#define HUV_NO_DIFFRACTION_ORDERS 6
#define HUV_PUPIL_SAMPLING_TOO_DENSE 7
#define HUV_UNSUPPORTED 8

const char* huv_last_error(void);
/* ABI version of this header (major * 100 + minor). */
int huv_abi_version(void);

/* ======================================================================== */
/* Host engine (FP64)                                                        */
/* ======================================================================== */

/* Optics: kind 0 = refractive / immersion projection optics
 * (optics/mod.rs:201-360), kind 1 = EUV mirror optics (optics/euv.rs:49-165). */
typedef struct {
  int kind;
  double na;
  double immersion_index;     /* kind 0 */
  double flare;               /* uniform flare fraction */
  double central_obscuration; /* kind 1, fraction of NA */
  double transmission;        /* kind 1, intensity */
  int paraxial_defocus;
} huv_optics;

/* Pupil fills (source.rs:1178-1248):
 *   0 conventional   p0 = sigma
 *   1 annular        p0 = sigma_inner, p1 = sigma_outer
 *   2 quadrupole     p0 = sigma_center, p1 = sigma_radius, p2 = opening_deg
 *   3 dipole         p0 = sigma_center, p1 = sigma_radius, p2 = orientation_deg
 *   4 gaussian       p0 = sigma
 *   5 sector poles   p0 = sigma_inner, p1 = sigma_outer, p2 = opening_deg,
 *                    p3 = first pole angle (deg), p4 = number of poles
 *                    (x-dipole: first 0, 2 poles; quasar: first 45, 4 poles) */
typedef struct {
  int kind;
  double p[6];
} huv_illumination;

/* Pupil fill intensity at (sx, sy) in sigma units. */
double huv_illumination_at(const huv_illumination* ill, double sx, double sy);

/* Adaptive (points_per_axis <= 0) or n x n midpoint source sampling
 * (aerial.rs:372-947, replaces `sample_source` aerial.rs:941).  Writes up to
 * `capacity` (sx, sy, weight) triples to xyw and the total count to *count
 * (call with capacity 0 to size the buffer). */
huv_status huv_sample_source(const huv_illumination* ill, int points_per_axis, double df_sigma,
                             double* xyw, int64_t capacity, int64_t* count);

/* Pupil function P(fx, fy; z, lambda) (optics/mod.rs:310-355, euv.rs:129-165). */
huv_status huv_pupil(const huv_optics* optics, double fx, double fy, double defocus_nm,
                     double wavelength_nm, double* re, double* im);

/* Kernel count selection (aerial.rs:992-1029). */
int huv_select_kernel_count(const double* values, int64_t nvals, double trace,
                            int64_t max_kernels, double fraction);

/* One SOCS kernel set (replaces KernelBuilder::build, aerial.rs:1041-1317).
 * method: -1 auto, 0 dense TCC, 1 dense Gram, 2 randomized. */
typedef struct huv_host_kernels huv_host_kernels;
typedef struct {
  int n;
  double pixel_nm;
  int num_kernels;
  int64_t num_support;
  int64_t num_rows;
  int method;
  int64_t num_columns;
  int support_exceeds_nyquist;
  double captured_energy_fraction;
  double clear_field;
  double pupil_peak;
  double normalization_divisor; /* clear field, or 1 for dark-field sets */
  double defocus_nm;
  double wavelength_nm;
} huv_kernels_info;

huv_status huv_kernels_build(const huv_optics* optics, const double* xyw, int64_t npoints, int n,
                             double pixel_nm, double defocus_nm, double wavelength_nm,
                             int max_kernels, double energy_fraction, int method,
                             huv_host_kernels** out);
huv_status huv_kernels_info_get(const huv_host_kernels* ks, huv_kernels_info* info);
/* Copy out (any pointer may be NULL): support (S flat indices i*n+j),
 * nonzero rows, absolute eigenvalues (K), values (K x S interleaved complex). */
huv_status huv_kernels_copy(const huv_host_kernels* ks, int64_t* support, int64_t* rows,
                            double* eigenvalues, double* values);
void huv_kernels_free(huv_host_kernels* ks);

/* Rasterise painted rectangles (later ones win) over a background with exact
 * area coverage (mask.rs:590, 1365-1571).  rects: nrects x (x, y, w, h, re, im);
 * out: n x n interleaved complex. */
huv_status huv_rasterize_rects(int n, double pixel_nm, int nrects, const double* rects,
                               double bg_re, double bg_im, double* out);

/* ======================================================================== */
/* Device backend (HIP)                                                      */
/* ======================================================================== */

typedef struct huv_ctx huv_ctx;
typedef struct huv_kset huv_kset;

/* Context on HIP device `device` with its own stream. */
huv_status huv_ctx_create(int device, huv_ctx** out);
void huv_ctx_destroy(huv_ctx* ctx);
/* Issue all further work on hip_stream (e.g. a framework's current stream;
 * NULL is the null stream).  huv_ctx_own_stream returns the context's own
 * (blocking) stream, the initial choice. */
huv_status huv_ctx_set_stream(huv_ctx* ctx, void* hip_stream);
void* huv_ctx_stream(huv_ctx* ctx);
void* huv_ctx_own_stream(huv_ctx* ctx);
huv_status huv_ctx_synchronize(huv_ctx* ctx);
/* Per-stage timing with HIP events recorded on the active stream around each
 * kernel launch (off by default).  Stages: 0 rows_band, 1 cols_band,
 * 2 socs_fields, 3 coarse_rows, 4 coarse_cols, 5 band_image, 6 ilt_pointwise,
 * 7 ilt_adjoint, 8 ilt_other, 9 rect_tables, 10 rect_band, 11 poly_band (9-11:
 * feature-list masks), 12 coarse (coarse_rows and coarse_cols fused in one
 * launch where the coarse grid allows it; 3 and 4 are then not used).
 * huv_ctx_stage_times synchronises, returns the
 * accumulated milliseconds and launch counts, and optionally resets them. */
huv_status huv_ctx_set_profiling(huv_ctx* ctx, int on);
huv_status huv_ctx_stage_times(huv_ctx* ctx, double* ms, int64_t* calls, int nstages, int reset);

/* Upload one kernel set (the sparse SparseKernelSet of aerial.rs:955-973):
 * support (S flat indices on the n grid), eigenvalues ALREADY divided by the
 * normalisation divisor (the KernelSet convention, aerial.rs:1912-1942),
 * values K x S interleaved complex, and the optics' flare fraction. */
huv_status huv_kset_upload(huv_ctx* ctx, int n, int num_kernels, int64_t num_support,
                           const int64_t* support, const double* eigenvalues,
                           const double* values, double flare, huv_kset** out);
/* Destroy a kernel set (before its context); drops graphs recorded with it. */
void huv_kset_destroy(huv_kset* ks);
/* Band half-width B and coarse grid m chosen for this set. */
huv_status huv_kset_shape(const huv_kset* ks, int* band, int* coarse);

/* Kernel sets of nplanes focus planes (defocus_nm[p]) built with the device's
 * help (replaces KernelBuilder::build, aerial.rs:1041-1317, for the dense TCC
 * and dense Gram paths): the TCC factor A[f, s] = sqrt(w_s) P(f/f_c + s; z,
 * lambda), its statistics (clear field, pupil peak, trace), its Gram product
 * (linalg.rs:181-232) and the kernels U = A V / sqrt(lambda) run on the GPU in
 * FP64; each plane's Hermitian eigendecomposition (linalg.rs:324-551) runs on
 * its own host thread.  Arguments as huv_kernels_build; out receives nplanes
 * sets (free each with huv_kernels_free).  A method that resolves to the
 * randomized path returns HUV_UNSUPPORTED (build it with huv_kernels_build).
 * Synchronises the context stream. */
huv_status huv_kernels_build_device(huv_ctx* ctx, const huv_optics* optics, const double* xyw,
                                    int64_t npoints, int n, double pixel_nm, int nplanes,
                                    const double* defocus_nm, double wavelength_nm,
                                    int max_kernels, double energy_fraction, int method,
                                    huv_host_kernels** out);

typedef struct {
  /* Multiply the mask spectrum by sinc(pi kx/n) sinc(pi ky/n) (mask.rs:804-815):
   * with a pixel-aligned raster this reproduces AerialImageEngine::compute(mask). */
  int pixel_transfer;
  /* Capture the call's launch sequence in a HIP graph the first time a given
   * argument set is seen (the call itself runs eagerly) and replay the graph
   * on later calls with the same device pointers, sizes and kernel sets.  The
   * buffers must stay allocated while the context may replay them. */
  int graph;
} huv_socs_opts;

/* Launch shape of the SOCS path for one (n, band, planes) combination:
 * coarse grid m (32 Q; sets the number of 32-point transform classes Q),
 * band-FFT factor n1 (four-step n = n1 x n / n1), and the field kernel's
 * x pass: xlanes = 1 or 2 lanes per band row (2 needs m >= 64), xnb = 4, 6
 * or 8 band columns per load block; chunks = 1..16 clip ranges issued one
 * after the other (each range's intermediates are then small enough to stay
 * in the device's last-level cache between stages; the default is sized for
 * that; results do not depend on it).  Zero fields select the built-in
 * defaults. */
typedef struct {
  int m;
  int n1;
  int xnb;
  int xlanes;
  int chunks;
} huv_socs_tuning;

/* Autotuner: when on, the first huv_socs_images / huv_socs_images_batch call
 * for a new (n, band, nplanes, max K) times every supported launch shape on
 * that call's own inputs (HIP events on the context stream, results of the
 * last run kept) and keeps the fastest for later calls.  The table can be
 * read and preset (e.g. persisted across processes). */
huv_status huv_ctx_set_autotune(huv_ctx* ctx, int on);
huv_status huv_ctx_tuning_get(huv_ctx* ctx, int n, int band, int nplanes, int kmax,
                              huv_socs_tuning* out, int* found);
huv_status huv_ctx_tuning_set(huv_ctx* ctx, int n, int band, int nplanes, int kmax,
                              const huv_socs_tuning* t);

/* Batched aerial images (replaces socs_intensity + normalize + apply_flare,
 * aerial.rs:1682-1759, for compute_from_transmittance / compute_through_focus).
 * masks: DEVICE pointer, nclips x n x n real transmittance (FP32).
 * planes: nplanes kernel sets sharing n, band and flare (focus planes).
 * out: DEVICE pointer, nclips x nplanes x n x n FP32, image (c, p) at
 * out + (c * nplanes + p) * n * n.  Asynchronous on the context stream. */
huv_status huv_socs_images(huv_ctx* ctx, int nclips, int n, const float* masks, int nplanes,
                           huv_kset* const* planes, float* out, const huv_socs_opts* opts);
/* huv_socs_images without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * intermediate buffers (row spectra, band boxes, coarse images, intensity
 * bands) are carved from a fixed array of HUV_NOALLOC_SCRATCH_MIB MiB per GPU
 * that the library holds until the process ends; a batch that needs more is
 * refused with HUV_INVALID_PARAMETER (pass fewer clips per call).  Set-up:
 * the transform tables and kernel stacks are not built here, so call
 * huv_socs_images once with the same kernel sets (any batch size) first;
 * otherwise HUV_INVALID_PARAMETER.  Uses the context's tuning-table entry for
 * the shape class if present (no autotuning) and records no graph, memset or
 * event.  All calls on one GPU share its scratch array: calls that overlap in
 * time on one GPU (two contexts, two streams) race and must be ordered by the
 * caller.  Asynchronous on the context stream. */
#define HUV_NOALLOC_SCRATCH_MIB 64
huv_status huv_socs_images_noalloc(huv_ctx* ctx, int nclips, int n, const float* masks,
                                   int nplanes, huv_kset* const* planes, float* out,
                                   int pixel_transfer);

/* Complex transmittance rasters (compute_from_transmittance /
 * compute_through_focus_from_transmittance with an Array2<Complex64> map,
 * aerial.rs:1789-1799,1829-1842): masks is a DEVICE pointer to nclips x n x n
 * interleaved complex FP32 (re, im) values; planes, out and opts as in
 * huv_socs_images (pixel_transfer and graph apply). */
huv_status huv_socs_images_complex(huv_ctx* ctx, int nclips, int n, const float* masks,
                                   int nplanes, huv_kset* const* planes, float* out,
                                   const huv_socs_opts* opts);
/* Images from mask spectra (compute_from_spectrum, aerial.rs:1805-1818):
 * spectra is a HOST pointer to nclips x n x n interleaved complex FP64 values
 * in Fft2D::forward's layout and scale (row ky, column kx, FFT index order,
 * unnormalised; fft2d.rs:11-13); only the |kx|, |ky| <= band bins are read
 * (rounded to FP32).  planes and out as in huv_socs_images.  Synchronises the
 * context stream before staging the band (pinned upload). */
huv_status huv_socs_images_spectrum(huv_ctx* ctx, int nclips, int n, const double* spectra,
                                    int nplanes, huv_kset* const* planes, float* out);

/* One group of equally sized clips of an uneven batch. */
typedef struct {
  int n;
  int nclips;
  const float* masks;      /* device, nclips x n x n */
  int nplanes;
  huv_kset* const* planes; /* nplanes kernel sets on the n grid */
  float* out;              /* device, nclips x nplanes x n x n */
} huv_clip_group;

/* Uneven batch: every group in order on the context stream; with
 * opts->graph the whole batch is one graph launch (batched
 * compute_through_focus over clips of several sizes). */
huv_status huv_socs_images_batch(huv_ctx* ctx, int ngroups, const huv_clip_group* groups,
                                 const huv_socs_opts* opts);
/* Same with host buffers (copies in, computes, copies out, synchronises;
 * opts->graph is ignored: the device copies are per call). */
huv_status huv_socs_images_host(huv_ctx* ctx, int nclips, int n, const float* masks,
                                int nplanes, huv_kset* const* planes, float* out,
                                const huv_socs_opts* opts);
/* Band spectrum stage alone (device pointers): M[ky][kx] = FFT_n(mask) for
 * |kx|, |ky| <= band; out: nclips x (2B+1) x (2B+1) interleaved complex FP32
 * with layout [kx+B][ky+B] (ky fastest). */
huv_status huv_band_spectrum(huv_ctx* ctx, int nclips, int n, int band, const float* masks,
                             float* out, int pixel_transfer);

/* Rectangle masks (host arrays).  Clip i is the rectangles
 * [offsets[i], offsets[i + 1]) of `rects`, each (x, y, w, h, re, im): centre
 * and size in nm at wafer scale, complex amplitude inside; later rectangles
 * win where they overlap (painter's rule); background[2 i], background[2 i + 1]
 * is the clip's amplitude elsewhere.  Geometry outside the field
 * [-n pixel_nm / 2, n pixel_nm / 2)^2 is clipped away.  A clip whose edge grid
 * exceeds 2^20 cells is rejected (rasterise it instead). */

/* Exact band spectrum of rectangle masks (replaces Mask::spectrum's analytic
 * path for rectangles, mask.rs:616-657 with compress / axis_coefficients /
 * add_term_spectrum at mask.rs:1365-1502): S[k] = n^2 c(m) e^{-i pi (mx+my)(n-1)/n}
 * for |kx|, |ky| <= band, FP64 evaluation; out: DEVICE, nclips x (2B+1) x (2B+1)
 * interleaved complex FP32 in huv_band_spectrum's layout [kx+B][ky+B].
 * Synchronises the context stream before staging the geometry. */
huv_status huv_band_spectrum_rects(huv_ctx* ctx, int nclips, int n, double pixel_nm, int band,
                                   const int64_t* offsets, const double* rects,
                                   const double* background, float* out);
/* Simple polygons of a batch (host arrays): clip i holds polygons
 * [poly_offsets[i], poly_offsets[i + 1]); polygon j has the vertices
 * [vertex_offsets[j], vertex_offsets[j + 1]) of `vertices` ((x, y) in nm,
 * either orientation) and complex amplitude values[2 j], values[2 j + 1].
 * A polygon is clipped to the field; one whose edges cross, or whose bounding
 * box meets another feature's of the same clip, is rejected (the reference
 * rasterises those, mask.rs:1190-1243; boxes that overlap by at most 1e-9 of
 * the field length count as disjoint, mask.rs:838-841,1099-1108), as is a
 * polygon of more than 8192 vertices or a call of more than 2^26 clipped
 * edges.  With polygons the band half-width is at most 511 (HUV_UNSUPPORTED
 * beyond). */
typedef struct {
  const int64_t* poly_offsets;
  const int64_t* vertex_offsets;
  const double* vertices;
  const double* values;
} huv_polygons;

/* Exact band spectrum of masks of rectangles and simple polygons (the
 * analytic path of Mask::spectrum, mask.rs:616-657, polygons by the edge sum
 * of mask.rs:1286-1352); arguments as in huv_band_spectrum_rects, polys may
 * be NULL. */
huv_status huv_band_spectrum_features(huv_ctx* ctx, int nclips, int n, double pixel_nm, int band,
                                      const int64_t* offsets, const double* rects,
                                      const huv_polygons* polys, const double* background,
                                      float* out);
/* Aerial images of masks of rectangles and simple polygons from their exact
 * spectrum (AerialImageEngine::compute(mask, defocus), aerial.rs:1775-1777). */
huv_status huv_socs_images_features(huv_ctx* ctx, int nclips, int n, double pixel_nm,
                                    const int64_t* offsets, const double* rects,
                                    const huv_polygons* polys, const double* background,
                                    int nplanes, huv_kset* const* planes, float* out,
                                    const huv_socs_opts* opts);
/* Aerial images of rectangle masks from their exact spectrum (replaces
 * AerialImageEngine::compute(mask, defocus), aerial.rs:1775-1777, for
 * rectangle masks: mask.spectrum then compute_from_spectrum).  planes and out
 * as in huv_socs_images; opts is not used (no pixel transfer applies to an
 * exact spectrum, and the per-call geometry upload is not captured in graphs). */
huv_status huv_socs_images_rects(huv_ctx* ctx, int nclips, int n, double pixel_nm,
                                 const int64_t* offsets, const double* rects,
                                 const double* background, int nplanes, huv_kset* const* planes,
                                 float* out, const huv_socs_opts* opts);

/* ======================================================================== */
/* ILT on the device (ilt.rs)                                                */
/* ======================================================================== */

typedef struct huv_ilt huv_ilt;

/* IltProblem settings (ilt.rs:159-224).  cost_kind 0 = image fidelity,
 * 1 = sigmoid resist contour (threshold, steepness).  beta = mask sigmoid
 * steepness; tv_weight / tv_epsilon / bin_weight = regularisers; absorber =
 * complex amplitude of the m = 0 state (0 binary, sqrt(T) e^{i phi} for an
 * attenuated PSM).  min_feature_nm > 0 adds the Gaussian density filter
 * m = sigma(beta G_s * theta), s = min_feature_nm / 1.349, on a grid of
 * pixel_nm pixels (ilt.rs:443-454, 524-535); its spectrum is applied on the
 * bins where it exceeds 1e-12 (HUV_UNSUPPORTED when that band is wider than
 * 127 bins).  gradient 0 = exact adjoint, 1 = the local proxy (the flare-adjoint
 * data gradient dC/dI in place of dC/dm, ilt.rs:719-721; steepest descent,
 * steps accepted on plain decrease, ilt.rs:789-801). */
typedef struct {
  int cost_kind;
  double threshold;
  double steepness;
  double beta;
  double tv_weight;
  double tv_epsilon;
  double bin_weight;
  double absorber_re;
  double absorber_im;
  double min_feature_nm;
  double pixel_nm;
  int gradient;
} huv_ilt_config;

/* Bind process conditions (one kernel set per focus, from huv_kset_upload
 * with eigenvalues divided by the clear field) with dose and weight
 * (replaces IltProblem::new, ilt.rs:389-478).  Coarse grids up to 256 (band
 * half-widths up to 63); conditions x targets x kernels per evaluation up to
 * 65535 on coarse grids above 64. */
huv_status huv_ilt_create(huv_ctx* ctx, int n, int nconds, huv_kset* const* conds,
                          const double* dose, const double* weight, const huv_ilt_config* cfg,
                          huv_ilt** out);
void huv_ilt_destroy(huv_ilt* ilt);

/* Batched cost (host, ntargets doubles) and, if grad != NULL, dC/dtheta
 * (device, ntargets x n x n) at theta (device) for targets (device)
 * (replaces IltProblem::cost / cost_and_gradient_with, ilt.rs:648-738). */
huv_status huv_ilt_cost_grad(huv_ilt* ilt, int ntargets, const float* targets, const float* theta,
                             double* cost, float* grad);

/* One CG (PR+) / steepest-descent stage with Armijo backtracking for every
 * target (replaces run_stage, ilt.rs:755-866).  theta (device) is updated in
 * place; cost_hist (host) receives ntargets x (max_iterations + 1) costs
 * (NaN after a target stops); iterations and termination (0 converged,
 * 1 max iterations, 2 line search failed) per target. */
huv_status huv_ilt_optimize(huv_ilt* ilt, int ntargets, const float* targets, float* theta,
                            int max_iterations, double learning_rate, double convergence_tol,
                            int use_cg, double* cost_hist, int* iterations, int* termination);
/* Same, ending each target as soon as its cost is <= stop_cost[t] (host,
 * ntargets values; termination 3): refinement of a warm start to a fixed
 * cost.  No counterpart in the reference; for the warm-started ILT. */
huv_status huv_ilt_optimize_to(huv_ilt* ilt, int ntargets, const float* targets, float* theta,
                               int max_iterations, double learning_rate, double convergence_tol,
                               int use_cg, const double* stop_cost, double* cost_hist,
                               int* iterations, int* termination);
/* run_stage with its trace (ilt.rs:755-866 with IltTrace, ilt.rs:745-750):
 * as huv_ilt_optimize_to with stop_cost nullable, plus gnorm_hist (host,
 * nullable, ntargets x (max_iterations + 1): the L2 norm of dC/dtheta at the
 * start and after each iteration, NaN after a target stops) and mask
 * snapshots: with snapshot_every > 0, target t's mask sigma(beta G_s * theta)
 * after its global iteration g = iteration_offset[t] + i is written to
 * snapshots (DEVICE, ntargets x nslots x n x n) slot g / snapshot_every - 1
 * whenever g is a multiple of snapshot_every and that slot is < nslots.
 * iteration_offset (host, nullable = zeros) holds the iterations each target
 * ran in earlier stages: max_iterations is then the whole budget, and target t
 * runs at most max_iterations - iteration_offset[t] iterations here. */
huv_status huv_ilt_run_stage(huv_ilt* ilt, int ntargets, const float* targets, float* theta,
                             int max_iterations, double learning_rate, double convergence_tol,
                             int use_cg, const double* stop_cost, double* cost_hist,
                             double* gnorm_hist, int snapshot_every, const int* iteration_offset,
                             int nslots, float* snapshots, int* iterations, int* termination);
/* Binarisation weight of later evaluations (the continuation of optimize_ilt,
 * ilt.rs:894-915, switches it between stages). */
huv_status huv_ilt_set_binarization(huv_ilt* ilt, double bin_weight);
/* m = sigma(beta G_s * theta) (IltProblem::mask, ilt.rs:537-542): theta and m
 * DEVICE, ntargets x n x n. */
huv_status huv_ilt_mask(huv_ilt* ilt, int ntargets, const float* theta, float* m);

/* ======================================================================== */
/* U-Net warm start for ILT                                                  */
/* ======================================================================== */

/* No counterpart in the reference (its ILT starts from the target,
 * ilt.rs:495-522): a 2-D U-Net mapping a target pattern to an initial theta.
 * `levels` resolution levels with base * 2^l channels, two 3x3 conv + ReLU per
 * encoder level and decoder stage, 2x2 max pooling, nearest 2x upsampling
 * with skip concatenation, 1x1 FP32 head.  BF16 weights and activations, FP32
 * accumulation on the matrix cores, FP32 master parameters and Adam.
 * Inputs and outputs are DEVICE FP32 [N][H][W]; H and W must be multiples of
 * 16 * 2^(levels-1).  Parameters are one flat FP32 vector (per conv: weights
 * [cout][3][3][cin] then bias [cout]; encoder levels, then decoder stages from
 * the deepest; head weights [base] then bias). */
typedef struct huv_unet huv_unet;
huv_status huv_unet_create(huv_ctx* ctx, int levels, int base_channels, uint64_t seed,
                           huv_unet** out);
void huv_unet_destroy(huv_unet* net);
huv_status huv_unet_num_params(const huv_unet* net, int64_t* count);
huv_status huv_unet_get_params(huv_unet* net, float* host);
huv_status huv_unet_set_params(huv_unet* net, const float* host);
huv_status huv_unet_forward(huv_unet* net, int n, int h, int w, const float* in, float* out);
/* Mean-squared error to `target` (host double) and its gradient, kept for
 * huv_unet_adam and copied to grad_host (nparams floats) when not NULL. */
huv_status huv_unet_loss_grad(huv_unet* net, int n, int h, int w, const float* in,
                              const float* target, double* loss, float* grad_host);
/* One Adam step with the last computed gradient (bias-corrected). */
huv_status huv_unet_adam(huv_unet* net, double lr, double beta1, double beta2, double eps);

#ifdef __cplusplus
}
#endif

#endif /* HUV_H */
