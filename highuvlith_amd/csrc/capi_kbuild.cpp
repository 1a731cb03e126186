// C ABI of the device-assisted SOCS kernel-set build (include/huv.h,
// huv_kernels_build_device): KernelBuilder::build (aerial.rs:1041-1317) for
// several focus planes, with the TCC factor A, its statistics, its Gram
// product (linalg.rs:181-232) and the kernels U = A V / sqrt(lambda) on the
// GPU in FP64, and the Hermitian eigendecomposition (linalg.rs:324-551) of
// every plane's Gram matrix on its own host thread.  Dense TCC and dense Gram
// paths; the randomized path stays with huv_kernels_build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/huv.h"
#include "capi_internal.hpp"
#include "hip/kbuild_kernels.hpp"

using namespace huvi;

namespace {

// Device buffers of one build call, freed on every exit path.
struct DevBuf {
  std::vector<void*> ptrs;
  ~DevBuf() {
    for (void* p : ptrs) hipFree(p);
  }
  template <class T>
  T* alloc(size_t count, hipError_t& e) {
    void* p = nullptr;
    e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return (T*)p;
  }
};

// What the device returns for one plane.
struct PlaneOut {
  std::vector<huv::cd> gram;  // D x D column-major
  std::vector<double> col_sum, col_dc, col_peak;
  std::vector<int> col_nz, row_nz;
  // host results
  std::vector<int64_t> support;
  std::vector<int> keep;      // kept columns (dense Gram) in order
  std::vector<double> vals;   // all eigenvalues, decreasing
  std::vector<huv::cd> vecs;  // eigenvectors, column-major
  int k = 0;
  double trace = 0, clear_field = 0, pupil_peak = 0;
  int64_t n_eff = 0;
  int status = huv::OK;
  std::string err;
};

}  // namespace

extern "C" {

huv_status huv_kernels_build_device(huv_ctx* c, const huv_optics* optics, const double* xyw,
                                    int64_t npoints, int n, double pixel_nm, int nplanes,
                                    const double* defocus_nm, double wavelength_nm,
                                    int max_kernels, double energy_fraction, int method,
                                    huv_host_kernels** out) {
  if (!c || !optics || !xyw || !defocus_nm || !out) return fail(HUV_INVALID_PARAMETER, "null argument");
  if (nplanes < 1) return fail(HUV_INVALID_PARAMETER, "need at least one focus plane");
  huv_status st = check_build_args(optics, n, pixel_nm, wavelength_nm, max_kernels,
                                   energy_fraction);
  if (st != HUV_OK) return st;
  for (int p = 0; p < nplanes; ++p)
    if (!std::isfinite(defocus_nm[p])) return fail(HUV_INVALID_PARAMETER, "defocus_nm: must be finite");
  try {
    std::vector<huv::SourcePoint> pts((size_t)npoints);
    for (int64_t i = 0; i < npoints; ++i) pts[i] = {xyw[3 * i], xyw[3 * i + 1], xyw[3 * i + 2]};
    const huv::Optics o = to_optics(optics);
    huv::Candidates cs;
    std::string err;
    int hs = huv::kernel_candidates(o, pts, n, pixel_nm, wavelength_nm, cs, err);
    if (hs != huv::OK) return fail(hs, err);
    huv::BuildSettings bs;
    bs.max_kernels = max_kernels;
    bs.kernel_energy_fraction = energy_fraction;
    bs.method = method;
    const int64_t n_cand = (int64_t)cs.idx.size();
    const int n_pts = (int)npoints;
    const int meth = huv::choose_build_method(n_cand, n_pts, bs);
    if (meth != huv::METHOD_DENSE_GRAM && meth != huv::METHOD_DENSE_TCC)
      return fail(HUV_UNSUPPORTED, "the device build covers the dense TCC and dense Gram paths; "
                                   "build randomized kernel sets with huv_kernels_build");
    const bool gram = meth == huv::METHOD_DENSE_GRAM;
    const int D = gram ? n_pts : (int)n_cand;
    const int64_t L = gram ? n_cand : n_pts;
    if (!gram && n_cand > (1 << 16)) return fail(HUV_UNSUPPORTED, "TCC too large for the device build");
    std::lock_guard<std::mutex> lk(c->mu);
    hipSetDevice(c->device);
    hipStream_t s = c->stream();
    hipError_t e;
    DevBuf db;
    double* d_fx = db.alloc<double>(n_cand, e);
    double* d_fy = d_fx ? db.alloc<double>(n_cand, e) : nullptr;
    double* d_pts = d_fy ? db.alloc<double>(3 * (size_t)n_pts, e) : nullptr;
    if (!d_pts) return hip_fail(e, "kernel build buffers");
    e = hipMemcpyAsync(d_fx, cs.fx.data(), n_cand * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_fy, cs.fy.data(), n_cand * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_pts, xyw, 3 * (size_t)n_pts * sizeof(double), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "kernel build upload");
    huvk::PupilParams pp;
    pp.euv = o.kind == huv::OPTICS_EUV;
    pp.paraxial = o.paraxial_defocus;
    pp.na = o.na;
    pp.n_medium = pp.euv ? 1.0 : o.immersion_index;
    pp.obscuration = pp.euv ? o.central_obscuration : 0.0;
    pp.amplitude = pp.euv ? std::sqrt(std::max(o.transmission, 0.0)) : 1.0;
    const int splits = huvk::gram_splits(D, L);
    double2* d_part = db.alloc<double2>((size_t)splits * D * D, e);
    if (!d_part) return hip_fail(e, "kernel build buffers");
    std::vector<double2*> d_A(nplanes), d_G(nplanes);
    std::vector<double*> d_stats(nplanes);
    std::vector<int*> d_flags(nplanes);
    std::vector<PlaneOut> po(nplanes);
    // device part of every plane: factor, statistics, Gram
    for (int p = 0; p < nplanes; ++p) {
      d_A[p] = db.alloc<double2>((size_t)n_cand * n_pts, e);
      d_G[p] = d_A[p] ? db.alloc<double2>((size_t)D * D, e) : nullptr;
      d_stats[p] = d_G[p] ? db.alloc<double>(3 * (size_t)n_pts, e) : nullptr;
      d_flags[p] = d_stats[p] ? db.alloc<int>((size_t)n_pts + n_cand, e) : nullptr;
      if (!d_flags[p]) return hip_fail(e, "kernel build buffers");
      e = huvk::launch_pupil_factor(d_fx, d_fy, n_cand, d_pts, n_pts, pp, defocus_nm[p],
                                    wavelength_nm, d_A[p], s);
      if (e == hipSuccess)
        e = huvk::launch_factor_stats(d_A[p], n_cand, d_pts, n_pts, d_stats[p],
                                      d_stats[p] + n_pts, d_stats[p] + 2 * n_pts, d_flags[p],
                                      gram ? nullptr : d_flags[p] + n_pts, s);
      if (e == hipSuccess)
        e = huvk::launch_gram(d_A[p], n_cand, gram ? 0 : 1, D, L, splits, d_part, d_G[p], s);
      if (e != hipSuccess) return hip_fail(e, "kernel build launch");
      PlaneOut& q = po[p];
      q.gram.resize((size_t)D * D);
      q.col_sum.resize(n_pts);
      q.col_dc.resize(n_pts);
      q.col_peak.resize(n_pts);
      q.col_nz.resize(n_pts);
      q.row_nz.resize(gram ? 0 : n_cand);
      e = hipMemcpyAsync(q.gram.data(), d_G[p], (size_t)D * D * sizeof(double2), hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipMemcpyAsync(q.col_sum.data(), d_stats[p], n_pts * sizeof(double), hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipMemcpyAsync(q.col_dc.data(), d_stats[p] + n_pts, n_pts * sizeof(double), hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipMemcpyAsync(q.col_peak.data(), d_stats[p] + 2 * n_pts, n_pts * sizeof(double), hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipMemcpyAsync(q.col_nz.data(), d_flags[p], n_pts * sizeof(int), hipMemcpyDeviceToHost, s);
      if (e == hipSuccess && !gram)
        e = hipMemcpyAsync(q.row_nz.data(), d_flags[p] + n_pts, n_cand * sizeof(int), hipMemcpyDeviceToHost, s);
      if (e != hipSuccess) return hip_fail(e, "kernel build copy");
    }
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "kernel build");
    // host part: compaction and the eigendecomposition, one thread per plane
    auto host_plane = [&](PlaneOut& q) {
      for (int sp = 0; sp < n_pts; ++sp) q.clear_field += q.col_dc[sp];
      for (int sp = 0; sp < n_pts; ++sp) q.pupil_peak = std::max(q.pupil_peak, q.col_peak[sp]);
      for (int sp = 0; sp < n_pts; ++sp)
        if (q.col_nz[sp]) {
          q.keep.push_back(sp);
          q.trace += q.col_sum[sp];
        }
      q.n_eff = (int64_t)q.keep.size();
      std::vector<int> idx;  // Gram rows / columns kept
      if (gram) {
        idx = q.keep;
        q.support = cs.idx;
      } else {
        for (int64_t r = 0; r < n_cand; ++r)
          if (q.row_nz[r]) {
            idx.push_back((int)r);
            q.support.push_back(cs.idx[r]);
          }
      }
      const int64_t m = (int64_t)idx.size();
      if (m == 0 || q.n_eff == 0) {
        q.status = huv::NO_DIFFRACTION_ORDERS;
        q.err = "no diffraction orders pass the pupil";
        return;
      }
      std::vector<huv::cd> h((size_t)m * m);
      for (int64_t b = 0; b < m; ++b)
        for (int64_t a = 0; a < m; ++a) h[b * m + a] = q.gram[(size_t)idx[b] * D + idx[a]];
      q.gram.clear();
      q.gram.shrink_to_fit();
      if (!huv::hermitian_eigen_desc(h, m, m, q.vals, q.vecs)) {
        q.status = huv::NUMERICAL_ERROR;
        q.err = "Hermitian eigensolver: QL iteration failed to converge";
        return;
      }
      q.k = huv::select_kernel_count(q.vals.data(), (int64_t)q.vals.size(), q.trace, max_kernels,
                                     energy_fraction);
    };
    std::vector<std::thread> th;
    for (int p = 0; p < nplanes; ++p) th.emplace_back(host_plane, std::ref(po[p]));
    for (auto& t : th) t.join();
    for (int p = 0; p < nplanes; ++p)
      if (po[p].status != huv::OK) return fail(po[p].status, po[p].err);
    // kernels: eigenvectors (dense TCC) or U = A V / sqrt(lambda) on the device (dense Gram)
    std::vector<std::vector<huv::cd>> values(nplanes);
    for (int p = 0; p < nplanes; ++p) {
      PlaneOut& q = po[p];
      const int k = q.k;
      if (k <= 0) return fail(HUV_NO_DIFFRACTION_ORDERS, "no diffraction orders pass the pupil");
      if (!gram) {
        values[p].assign(q.vecs.begin(), q.vecs.begin() + (size_t)k * q.support.size());
        continue;
      }
      const int ne = (int)q.n_eff;
      std::vector<double> is(k);
      for (int cc = 0; cc < k; ++cc) is[cc] = 1.0 / std::sqrt(q.vals[cc]);
      double2* d_V = db.alloc<double2>((size_t)k * ne, e);
      double* d_is = d_V ? db.alloc<double>(k, e) : nullptr;
      int* d_keep = d_is ? db.alloc<int>(ne, e) : nullptr;
      double2* d_U = d_keep ? db.alloc<double2>((size_t)k * n_cand, e) : nullptr;
      if (!d_U) return hip_fail(e, "kernel build buffers");
      e = hipMemcpyAsync(d_V, q.vecs.data(), (size_t)k * ne * sizeof(double2), hipMemcpyHostToDevice, s);
      if (e == hipSuccess) e = hipMemcpyAsync(d_is, is.data(), k * sizeof(double), hipMemcpyHostToDevice, s);
      if (e == hipSuccess) e = hipMemcpyAsync(d_keep, q.keep.data(), ne * sizeof(int), hipMemcpyHostToDevice, s);
      if (e == hipSuccess)
        e = huvk::launch_kernels_from_vectors(d_A[p], n_cand, d_V, d_is, d_keep, ne, k, d_U, s);
      values[p].resize((size_t)k * n_cand);
      if (e == hipSuccess)
        e = hipMemcpyAsync(values[p].data(), d_U, (size_t)k * n_cand * sizeof(double2),
                           hipMemcpyDeviceToHost, s);
      if (e != hipSuccess) return hip_fail(e, "kernel build kernels");
    }
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "kernel build");
    std::vector<std::unique_ptr<huv_host_kernels>> res(nplanes);
    for (int p = 0; p < nplanes; ++p) {
      PlaneOut& q = po[p];
      res[p] = std::make_unique<huv_host_kernels>();
      std::vector<double> eig(q.vals.begin(), q.vals.begin() + q.k);
      hs = huv::assemble_kernel_set(n, pixel_nm, defocus_nm[p], wavelength_nm, q.support,
                                    std::move(eig), std::move(values[p]), q.trace, q.clear_field,
                                    q.pupil_peak, meth, q.n_eff, cs.exceeds_nyquist, res[p]->ks,
                                    err);
      if (hs != huv::OK) return fail(hs, err);
    }
    for (int p = 0; p < nplanes; ++p) out[p] = res[p].release();
    return HUV_OK;
  } catch (const std::exception& ex) {
    return fail(HUV_INTERNAL_ERROR, ex.what());
  }
}

}  // extern "C"
