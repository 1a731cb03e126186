// Internal state behind the opaque handles of include/huv.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/huv.h"
#include "hip/socs_kernels.hpp"
#include "host/engine.hpp"

// A kernel set built on the host (huv_kernels_build) or with the device's
// help (huv_kernels_build_device).
struct huv_host_kernels {
  huv::KernelSet ks;
};

namespace huvi {

huv::Optics to_optics(const huv_optics* o);
// The argument checks huv_kernels_build and huv_kernels_build_device share.
huv_status check_build_args(const huv_optics* optics, int n, double pixel_nm,
                            double wavelength_nm, int max_kernels, double energy_fraction);

huv_status fail(huv_status code, const std::string& msg);
huv_status hip_fail(hipError_t e, const char* what);

// Device scratch owned by a context, grown on demand and reused across calls.
struct Workspace {
  enum Slot { TMP_ROWS = 0, BAND, BAND_PAIRED, COARSE, TMP_COARSE, IBAND, ILT0, ILT1, ILT2, ILT3,
              ILT4, ILT5, ILT6, ILT7, ILT8, ILT9, ILT10, ILT11, RECT_TAB, kSlots };
  void* ptr[kSlots] = {};
  size_t size[kSlots] = {};
  uint64_t generation = 0;  // bumped whenever a slot is reallocated
  void* get(int slot, size_t bytes, hipError_t& err);
  void release();
};

struct Ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t active = nullptr;  // stream all work is issued on (may be the null stream)
  std::mutex mu;
  Workspace ws;
  std::map<int, float2*> tw;
  std::map<long, float*> sinc;
  // Device copies of per-plane pointer tables and packed kernel stacks, keyed
  // by the ids of the kernel sets in a plane list (dropped when a set dies).
  std::map<std::vector<uint64_t>, void*> plane_tables;
  struct PlanePack {
    float2* kern = nullptr;  // [Q][nplanes][Kmax] boxes (W x W or column-paired)
    float* lam = nullptr;    // [nplanes][Kmax], zero-padded
    int Kmax = 0;
  };
  // keyed by the kernel-set ids followed by the box half-width
  std::map<std::vector<uint64_t>, PlanePack> plane_packs;
  // Launch shapes chosen by the autotuner, keyed by (n, band, nplanes, Kmax).
  std::map<std::vector<int>, huv_socs_tuning> tuned;
  bool autotune = false;
  // Captured launch sequences, keyed by the call's pointers, sizes and
  // kernel-set ids; valid while the workspace generation is unchanged.
  struct GraphEntry {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t generation = 0;
    std::vector<uint64_t> ksets;
  };
  std::map<std::vector<uint64_t>, GraphEntry> graphs;
  // Pinned host staging area and its device mirror for small per-step
  // uploads (the ILT driver's coefficients and row lists), grown on demand.
  char* pin_h = nullptr;
  char* pin_d = nullptr;
  size_t pin_cap = 0;
  bool staging(size_t bytes, hipError_t& e);
  // Pinned, device-mapped host buffer for small per-step downloads (the ILT
  // driver's reduction and cost partials, which its kernels write there
  // directly through dl_d), grown on demand.
  char* dl_h = nullptr;
  char* dl_d = nullptr;
  size_t dl_cap = 0;
  bool download_area(size_t bytes, hipError_t& e);
  void drop_graph(GraphEntry& g);
  void forget_kset(uint64_t id);
  // Optional per-stage timing with HIP events on the active stream.
  bool profiling = false;
  struct StageEvents { int stage; hipEvent_t a, b; };
  std::vector<StageEvents> pending;
  double stage_ms[16] = {};
  int64_t stage_calls[16] = {};
  void stage_begin(int stage, hipEvent_t* a);
  void stage_end(int stage, hipEvent_t a);
  void collect_stage_times();
  hipStream_t stream() const { return active; }
  const float2* twiddles(int N, hipError_t& err);
  const float* sinc_table(int n, int B, bool pixel_transfer, hipError_t& err);
  // This device's fixed scratch array (resolved once per device at context
  // creation) and, while a no-allocation call runs, the part of it handed out
  // so far: workspace requests are then carved from it in order instead of
  // growing the workspace.
  char* fixed = nullptr;
  size_t fixed_cap = 0;
  bool fixed_mode = false;
  size_t fixed_used = 0;
  void* buf(int slot, size_t bytes, hipError_t& err);
  ~Ctx();
};

// Device bytes of one forward call's workspace slots (socs_forward), each
// rounded up to 256 bytes: band-FFT row spectra, band box, its column-paired
// copy (two x lanes), S coarse-image partials, coarse row band, intensity band.
struct SocsSizes {
  size_t tmp_rows = 0, band = 0, band_paired = 0, coarse = 0, tmp_coarse = 0, iband = 0;
  size_t total() const { return tmp_rows + band + band_paired + coarse + tmp_coarse + iband; }
};
SocsSizes socs_sizes(int nclips, int n, int B, int m, bool paired, int nplanes, int S,
                     bool raster);

enum Stage { ST_ROWS_BAND = 0, ST_COLS_BAND, ST_FIELDS, ST_COARSE_ROWS, ST_COARSE_COLS,
             ST_BAND_IMAGE, ST_ILT_POINTWISE, ST_ILT_ADJOINT, ST_ILT_OTHER, ST_RECT_TABLES,
             ST_RECT_BAND, ST_POLY_BAND, ST_COARSE, ST_COUNT };

// Band spectrum of nclips real n x n masks: bins |k| <= B computed, stored
// in boxes of half-width BB >= B (zero beyond B).  Layout Mq[clip][kx + BB][ky + BB];
// Mp (optional) receives the column-paired copy (huvk::launch_cols_band).
huv_status band_stage(Ctx* c, int nclips, int n, int B, int BB, const float* masks, float2* Mq,
                      bool pixel_transfer, int n1 = 0, float2* Mp = nullptr);

// Rectangle masks of a batch (host arrays): clip i holds rectangles
// [offsets[i], offsets[i + 1]) of rects (x, y, w, h, re, im; centre and size
// in nm, later ones win) over background (re, im) background[2 i ..].
struct RectBatch {
  double pixel_nm;
  const int64_t* offsets;
  const double* rects;
  const double* background;
  const huv_polygons* polys = nullptr;  // optional simple polygons per clip
};
// Exact band spectrum of rectangle masks into Mq (and Mp), same layout as band_stage.
huv_status rect_band_stage(Ctx* c, int nclips, int n, int B, int BB, const RectBatch& rb,
                           float2* Mq, float2* Mp);
huv_status validate_rects(int nclips, int n, const RectBatch& rb);

// Band spectrum of nclips complex n x n rasters (device, interleaved FP32):
// same output as band_stage, every row and both signs of kx transformed.
huv_status band_stage_complex(Ctx* c, int nclips, int n, int B, int BB, const float2* masks,
                              float2* Mq, bool pixel_transfer, int n1 = 0, float2* Mp = nullptr);
// Band of nclips host spectra (n x n interleaved complex FP64, the FFT layout
// of Fft2D::forward) into Mq (and Mp): the |k| <= B bins, rounded to FP32 and
// uploaded through the pinned staging area (synchronises the stream first).
huv_status spectrum_band_stage(Ctx* c, int nclips, int n, int B, int BB, const double* spectra,
                               float2* Mq, float2* Mp);

// What a forward call images: real or complex device rasters, feature lists
// (exact spectrum), or host spectra.
struct MaskSource {
  enum Kind { REAL = 0, COMPLEX, FEATURES, SPECTRUM };
  Kind kind = REAL;
  const float* raster = nullptr;     // REAL: [nclips][n][n]; COMPLEX: interleaved pairs
  const RectBatch* rects = nullptr;  // FEATURES
  const double* spectra = nullptr;   // SPECTRUM (host)
  static MaskSource real(const float* p) { MaskSource s; s.raster = p; return s; }
};

huv_status validate_planes(Ctx* c, int n, int nplanes, huv_kset* const* planes);
huv_status plane_table(Ctx* c, int nplanes, huv_kset* const* planes, const int** dK,
                       const float2* const** dkern, const float* const** dlam);
// Per-class pre-twisted kernel stacks at coarse size m (column-paired boxes
// when `paired`, for the two-lane field kernel).
huv_status plane_pack(Ctx* c, int nplanes, huv_kset* const* planes, int m, bool paired,
                      const float2** kern, const float** lam, int* Kmax);
// Forward images; tune = NULL uses the kernel sets' own coarse grid and the
// default launch shapes (the ILT adjoint relies on that band layout).
// pixel_transfer applies to rasters only.
// split_batch > 0 chooses the field kernel's split of the kernel sum as for a
// batch of that many clips, whatever nclips is: images then do not depend on
// which other clips share the call (the ILT relies on it).  fid: the last
// stage evaluates the ILT fidelity of each image instead of writing it (out
// unused).
huv_status socs_forward(Ctx* c, int nclips, int n, const MaskSource& src, int nplanes,
                        huv_kset* const* planes, float* out, bool pixel_transfer, float2** Mq_out,
                        const huv_socs_tuning* tune = nullptr, int max_splits = 0,
                        int split_batch = 0, const huvk::FidelityEpi* fid = nullptr);

}  // namespace huvi

struct huv_ctx : huvi::Ctx {};

struct huv_kset {
  huv_ctx* ctx = nullptr;
  uint64_t id = 0;  // unique per process, keys the plane caches
  int n = 0, K = 0, B = 0, m = 0;
  int BB = 0;       // stored box half-width (coarse_band(m) >= B)
  int64_t S = 0;
  double flare = 0.0;
  float2* kern = nullptr;  // K x W x W (W = 2 BB + 1), [k][kx+BB][ky+BB]
  float* lam = nullptr;    // K, eigenvalue / n^4
  std::vector<double> lam_host;
  std::vector<int64_t> support;
  std::vector<double> values;
  ~huv_kset() {
    if (kern) hipFree(kern);
    if (lam) hipFree(lam);
  }
};
