// SOCS kernel-set construction on the host in FP64
// (crates/highuvlith-core/src/aerial.rs:992-1317): candidate mask
// frequencies |f| <= (1 + sigma_max) NA/lambda, factor A[f, s] =
// sqrt(w_s) P(f/f_c + s; z, lambda), and the leading eigenpairs of
// TCC = A A^H by the dense TCC, dense Gram or randomized path.  Also the
// exact-coverage rasteriser of painted rectangles (mask.rs:1365-1571).
#include <algorithm>
#include <cmath>
#include <map>

#include "engine.hpp"

namespace huv {

namespace {
const size_t MAX_FACTOR_BYTES = size_t(1) << 30;
const int64_t DENSE_FAST_DIM = 320;
const int64_t DENSE_MAX_DIM = 1536;
const int64_t RANDOMIZED_OVERSAMPLE = 16;
const int RANDOMIZED_POWER_ITERS = 3;
const uint64_t RANDOMIZED_SEED = 0x5EED50C5ull;
const double EIGEN_ZERO_REL = 1e-12;
const double DEGENERACY_REL = 1e-6;
const double DARK_FIELD_FRACTION = 1e-2;

double signed_index(int64_t k, int64_t n) { return k < n / 2 ? (double)k : (double)(k - n); }
}  // namespace

double KernelSet::normalization_divisor() const {
  const bool bright = clear_field > 0.0 && clear_field >= DARK_FIELD_FRACTION * pupil_peak;
  return bright ? clear_field : 1.0;
}

int select_kernel_count(const double* values, int64_t nvals, double trace,
                        int64_t max_kernels, double fraction) {
  const double lead = nvals > 0 ? values[0] : 0.0;
  if (std::isnan(lead) || lead <= 0.0) return 0;
  const double tol = EIGEN_ZERO_REL * lead;
  int64_t n_valid = 0;
  while (n_valid < nvals && values[n_valid] > tol) ++n_valid;
  const int64_t cap = std::min(max_kernels, n_valid);
  int64_t k = cap;
  if (fraction < 1.0) {
    double acc = 0.0;
    for (int64_t i = 0; i < cap; ++i) {
      acc += values[i];
      if (acc >= fraction * trace) {
        k = i + 1;
        break;
      }
    }
  }
  auto same = [](double a, double b) {
    return std::fabs(a - b) <= DEGENERACY_REL * std::max(std::fabs(a), std::fabs(b));
  };
  if (k > 0 && k < n_valid && same(values[k - 1], values[k])) {
    int64_t end = k;
    while (end < n_valid && same(values[end - 1], values[end])) ++end;
    if (end <= max_kernels) {
      k = end;
    } else {
      while (k > 0 && same(values[k - 1], values[k])) --k;
      if (k == 0) k = end;
    }
  }
  return (int)k;
}

int kernel_candidates(const Optics& optics, const std::vector<SourcePoint>& points, int n,
                      double pixel_nm, double wavelength_nm, Candidates& out, std::string& err) {
  const double df = 1.0 / (n * pixel_nm);
  const double cutoff = optics.na / wavelength_nm;
  if (!(std::isfinite(cutoff) && cutoff > 0.0)) {
    err = "cutoff_frequency: optics cutoff NA/lambda must be positive and finite";
    return INVALID_PARAMETER;
  }
  double sigma_max = 0.0;
  for (const auto& p : points) sigma_max = std::max(sigma_max, std::hypot(p.sx, p.sy));
  const double radius = cutoff * (1.0 + sigma_max);
  const double r2 = std::pow(radius / df, 2) * (1.0 + 1e-12);
  out = Candidates();
  out.exceeds_nyquist = radius > (double)(n / 2) * df;
  for (int64_t i = 0; i < n; ++i) {
    const double ky = signed_index(i, n);
    if (ky * ky > r2) continue;
    for (int64_t j = 0; j < n; ++j) {
      const double kx = signed_index(j, n);
      if (kx * kx + ky * ky <= r2) {
        out.idx.push_back(i * n + j);
        out.fx.push_back(kx * df / cutoff);
        out.fy.push_back(ky * df / cutoff);
      }
    }
  }
  const int64_t n_cand = (int64_t)out.idx.size(), n_cols = (int64_t)points.size();
  if (n_cand == 0 || n_cols == 0) {
    err = "no diffraction orders pass the pupil";
    return NO_DIFFRACTION_ORDERS;
  }
  if ((size_t)n_cand * (size_t)n_cols * sizeof(cd) > MAX_FACTOR_BYTES) {
    err = "pupil sampling too dense: " + std::to_string(n_cand) + " samples, max " +
          std::to_string(MAX_FACTOR_BYTES / (sizeof(cd) * n_cols));
    return PUPIL_SAMPLING_TOO_DENSE;
  }
  return OK;
}

int choose_build_method(int64_t n_cand, int64_t n_cols, const BuildSettings& settings) {
  const int64_t min_dim = std::min(n_cand, n_cols);
  const int64_t subspace = settings.max_kernels + RANDOMIZED_OVERSAMPLE;
  const bool dense = min_dim <= DENSE_FAST_DIM || (min_dim <= DENSE_MAX_DIM && subspace * 3 >= min_dim);
  if (settings.method != METHOD_AUTO) return settings.method;
  return !dense ? METHOD_RANDOMIZED : (n_cand <= n_cols ? METHOD_DENSE_TCC : METHOD_DENSE_GRAM);
}

int build_kernel_set(const Optics& optics, const std::vector<SourcePoint>& points, int n,
                     double pixel_nm, double defocus_nm, double wavelength_nm,
                     const BuildSettings& settings, KernelSet& out, std::string& err) {
  Candidates cs;
  const int cst = kernel_candidates(optics, points, n, pixel_nm, wavelength_nm, cs, err);
  if (cst != OK) return cst;
  struct Cand { int64_t idx; double fx, fy; };
  std::vector<Cand> cand(cs.idx.size());
  for (size_t i = 0; i < cand.size(); ++i) cand[i] = {cs.idx[i], cs.fx[i], cs.fy[i]};
  const bool exceeds_nyquist = cs.exceeds_nyquist;
  const int64_t n_cand = (int64_t)cand.size();
  const int64_t n_pts = (int64_t)points.size();
  const int64_t n_cols = n_pts;
  auto eval = [&](const Cand& c, const SourcePoint& p) {
    return pupil_function(optics, c.fx + p.sx, c.fy + p.sy, defocus_nm, wavelength_nm) *
           std::sqrt(p.weight);
  };
  const int64_t max_k = settings.max_kernels;
  const double fraction = settings.kernel_energy_fraction;
  const int64_t subspace = max_k + RANDOMIZED_OVERSAMPLE;
  const int method = choose_build_method(n_cand, n_cols, settings);

  std::vector<int64_t> support;
  std::vector<double> eig;
  std::vector<cd> values;  // K x len(support)
  double trace = 0, clear_field = 0, pupil_peak = 0;
  int64_t n_eff = 0;

  if (method == METHOD_DENSE_TCC) {
    // Rows of conj(A): TCC = Gram of these vectors.
    std::vector<cd> rows((size_t)(n_cand * n_cols));
    for (int64_t r = 0; r < n_cand; ++r)
      for (int64_t s = 0; s < n_pts; ++s) rows[r * n_cols + s] = std::conj(eval(cand[r], points[s]));
    for (int64_t s = 0; s < n_cols; ++s) clear_field += std::norm(rows[s]);
    for (int64_t b = 0; b < n_cand * n_cols; ++b)
      pupil_peak = std::max(pupil_peak, std::norm(rows[b]) / points[b % n_pts].weight);
    std::vector<char> col_nz((size_t)n_cols, 0);
    for (int64_t r = 0; r < n_cand; ++r)
      for (int64_t c = 0; c < n_cols; ++c)
        if (rows[r * n_cols + c] != cd(0, 0)) col_nz[c] = 1;
    std::vector<int64_t> keep;
    for (int64_t c = 0; c < n_cols; ++c)
      if (col_nz[c]) keep.push_back(c);
    n_eff = (int64_t)keep.size();
    ColMatrix x;
    x.rows = n_eff;
    for (int64_t r = 0; r < n_cand; ++r) {
      bool any = false;
      for (int64_t c : keep) any = any || rows[r * n_cols + c] != cd(0, 0);
      if (!any) continue;
      for (int64_t c : keep) x.data.push_back(rows[r * n_cols + c]);
      support.push_back(cand[r].idx);
    }
    x.cols = (int64_t)support.size();
    if (x.cols == 0 || n_eff == 0) {
      err = "no diffraction orders pass the pupil";
      return NO_DIFFRACTION_ORDERS;
    }
    trace = x.frobenius_norm_sq();
    std::vector<double> vals;
    std::vector<cd> vecs;
    if (!hermitian_eigen_desc(x.gram(), x.cols, x.cols, vals, vecs)) {
      err = "Hermitian eigensolver: QL iteration failed to converge";
      return NUMERICAL_ERROR;
    }
    const int k = select_kernel_count(vals.data(), (int64_t)vals.size(), trace, max_k, fraction);
    eig.assign(vals.begin(), vals.begin() + k);
    values.assign(vecs.begin(), vecs.begin() + (size_t)k * x.cols);
  } else {
    // Columns of A, source point by source point.
    ColMatrix a;
    a.rows = n_cand;
    a.cols = n_cols;
    a.data.resize((size_t)(n_cand * n_cols));
    for (int64_t s = 0; s < n_pts; ++s)
      for (int64_t r = 0; r < n_cand; ++r) a.data[s * n_cand + r] = eval(cand[r], points[s]);
    for (int64_t s = 0; s < n_cols; ++s) clear_field += std::norm(a.data[s * n_cand]);
    for (int64_t s = 0; s < n_pts; ++s) {
      double pk = 0.0;
      for (int64_t r = 0; r < n_cand; ++r)
        pk = std::max(pk, std::norm(a.data[s * n_cand + r]) / points[s].weight);
      pupil_peak = std::max(pupil_peak, pk);
    }
    int64_t ne = 0;
    for (int64_t c = 0; c < n_cols; ++c) {
      bool nz = false;
      for (int64_t r = 0; r < n_cand && !nz; ++r) nz = a.data[c * n_cand + r] != cd(0, 0);
      if (nz) {
        if (ne != c)
          std::copy(a.data.begin() + c * n_cand, a.data.begin() + (c + 1) * n_cand,
                    a.data.begin() + ne * n_cand);
        ++ne;
      }
    }
    if (ne == 0) {
      err = "no diffraction orders pass the pupil";
      return NO_DIFFRACTION_ORDERS;
    }
    n_eff = ne;
    a.data.resize((size_t)(ne * n_cand));
    a.cols = ne;
    trace = a.frobenius_norm_sq();
    for (auto& c : cand) support.push_back(c.idx);
    if (method == METHOD_DENSE_GRAM) {
      std::vector<double> vals;
      std::vector<cd> v;
      if (!hermitian_eigen_desc(a.gram(), ne, ne, vals, v)) {
        err = "Hermitian eigensolver: QL iteration failed to converge";
        return NUMERICAL_ERROR;
      }
      const int k = select_kernel_count(vals.data(), (int64_t)vals.size(), trace, max_k, fraction);
      eig.assign(vals.begin(), vals.begin() + k);
      values.assign((size_t)k * n_cand, cd(0, 0));
      // U = A V / sqrt(lambda)
      for (int c = 0; c < k; ++c) {
        cd* u = values.data() + (size_t)c * n_cand;
        const double is = 1.0 / std::sqrt(vals[c]);
        for (int64_t s = 0; s < ne; ++s) {
          const cd b = v[c * ne + s] * is;
          if (b == cd(0, 0)) continue;
          const cd* acol = a.col(s);
          for (int64_t r = 0; r < n_cand; ++r) u[r] += b * acol[r];
        }
      }
    } else {
      const int64_t l = std::min(subspace, std::min(n_cand, ne));
      std::vector<double> vals;
      ColMatrix u;
      randomized_eigen(a, l, RANDOMIZED_POWER_ITERS, RANDOMIZED_SEED, vals, u);
      const int k =
          select_kernel_count(vals.data(), (int64_t)vals.size(), trace, std::min(max_k, l), fraction);
      eig.assign(vals.begin(), vals.begin() + k);
      values.assign(u.data.begin(), u.data.begin() + (size_t)k * n_cand);
    }
  }
  return assemble_kernel_set(n, pixel_nm, defocus_nm, wavelength_nm, std::move(support),
                             std::move(eig), std::move(values), trace, clear_field, pupil_peak,
                             method, n_eff, exceeds_nyquist, out, err);
}

int assemble_kernel_set(int n, double pixel_nm, double defocus_nm, double wavelength_nm,
                        std::vector<int64_t> support, std::vector<double> eig,
                        std::vector<cd> values, double trace, double clear_field,
                        double pupil_peak, int method, int64_t n_eff, bool exceeds_nyquist,
                        KernelSet& out, std::string& err) {
  if (eig.empty() || std::isnan(trace) || trace <= 0.0) {
    err = "no diffraction orders pass the pupil";
    return NO_DIFFRACTION_ORDERS;
  }
  double captured = 0.0;
  for (double v : eig) captured += v;
  captured /= trace;
  std::vector<int64_t> rows_present;
  for (int64_t idx : support) {
    const int64_t r = idx / n;
    if (rows_present.empty() || rows_present.back() != r) rows_present.push_back(r);
  }
  out = KernelSet();
  out.n = n;
  out.pixel_nm = pixel_nm;
  out.support = std::move(support);
  out.nonzero_rows = std::move(rows_present);
  out.eigenvalues = std::move(eig);
  out.values = std::move(values);
  out.captured_energy_fraction = captured;
  out.clear_field = clear_field;
  out.pupil_peak = pupil_peak;
  out.defocus_nm = defocus_nm;
  out.wavelength_nm = wavelength_nm;
  out.method = method;
  out.num_columns = n_eff;
  out.support_exceeds_nyquist = exceeds_nyquist;
  return OK;
}

// ---------------------------------------------------------------------------
// Rasteriser: painted rectangles over a background, resolved into disjoint
// cells (later features win), each cell added with its exact per-pixel area
// coverage (separable x and y coverage).
// ---------------------------------------------------------------------------
namespace {
void interval_coverage(std::vector<double>& cov, double a, double b, int n, double pixel,
                       double half) {
  const double x0 = -half;
  const double lo = std::max(a, x0), hi = std::min(b, half);
  if (hi <= lo) return;
  int64_t ja = std::min<int64_t>(n - 1, (int64_t)std::max(0.0, std::floor((lo - x0) / pixel)));
  int64_t jb = std::min<int64_t>(n - 1, (int64_t)std::max(0.0, std::ceil((hi - x0) / pixel) - 1.0));
  for (int64_t j = ja; j <= jb; ++j) {
    const double left = x0 + j * pixel, right = x0 + (j + 1) * pixel;
    if (lo <= left && hi >= right) {
      cov[j] += 1.0;
    } else {
      const double o = std::min(hi, right) - std::max(lo, left);
      if (o > 0.0) cov[j] += o / pixel;
    }
  }
}

void dedup(std::vector<double>& e, double eps) {
  std::sort(e.begin(), e.end());
  std::vector<double> o;
  for (double v : e)
    if (o.empty() || v - o.back() > eps) o.push_back(v);
  e.swap(o);
}

// First edge >= v - eps (partition point).
int64_t edge_index(const std::vector<double>& e, double v, double eps) {
  return (int64_t)(std::lower_bound(e.begin(), e.end(), v - eps) - e.begin());
}
}  // namespace

void rasterize_rects(const std::vector<RectFeature>& rects, cd bg, int n, double pixel,
                     std::vector<cd>& out) {
  const double half = n * pixel / 2.0;
  const double eps = 1e-12 * (2.0 * half);
  out.assign((size_t)n * n, bg);
  struct R { double x0, x1, y0, y1; cd v; };
  std::vector<R> rs;
  for (const auto& f : rects) {
    if (!(f.w > 0 && f.h > 0)) continue;
    const double x0 = std::max(f.x - f.w / 2, -half), x1 = std::min(f.x + f.w / 2, half);
    const double y0 = std::max(f.y - f.h / 2, -half), y1 = std::min(f.y + f.h / 2, half);
    if (x1 - x0 <= 0 || y1 - y0 <= 0) continue;
    rs.push_back({x0, x1, y0, y1, f.value});
  }
  if (rs.empty()) return;
  std::vector<double> xe{-half, half}, ye{-half, half};
  for (auto& r : rs) {
    xe.push_back(r.x0); xe.push_back(r.x1);
    ye.push_back(r.y0); ye.push_back(r.y1);
  }
  dedup(xe, eps);
  dedup(ye, eps);
  const int64_t nx = (int64_t)xe.size() - 1, ny = (int64_t)ye.size() - 1;
  std::vector<cd> cells((size_t)(nx * ny), bg);
  for (auto& r : rs) {
    const int64_t i0 = edge_index(ye, r.y0, eps), i1 = edge_index(ye, r.y1, eps);
    const int64_t j0 = edge_index(xe, r.x0, eps), j1 = edge_index(xe, r.x1, eps);
    for (int64_t i = i0; i < std::min(i1, ny); ++i)
      for (int64_t j = j0; j < std::min(j1, nx); ++j) cells[i * nx + j] = r.v;
  }
  std::vector<double> covx((size_t)n), covy((size_t)n);
  for (int64_t i = 0; i < ny; ++i) {
    std::fill(covy.begin(), covy.end(), 0.0);
    interval_coverage(covy, ye[i], ye[i + 1], n, pixel, half);
    int64_t j = 0;
    while (j < nx) {
      const cd v = cells[i * nx + j];
      int64_t e = j + 1;
      while (e < nx && cells[i * nx + e] == v) ++e;
      if (v != bg) {
        std::fill(covx.begin(), covx.end(), 0.0);
        interval_coverage(covx, xe[j], xe[e], n, pixel, half);
        const cd w = v - bg;
        for (int64_t py = 0; py < n; ++py) {
          if (covy[py] == 0.0) continue;
          for (int64_t px = 0; px < n; ++px)
            if (covx[px] != 0.0) out[py * n + px] += w * (covx[px] * covy[py]);
        }
      }
      j = e;
    }
  }
}

// Disjoint cells of painted rectangles (mask.rs:1365-1424, compress): the
// rectangles are clipped to the field, the field and rectangle edges are
// merged within 1e-12 L, each rectangle paints its cells in order (the later
// one wins), runs of equal value along x become one cell and consecutive
// rows with identical content one row.  Cells equal to the background are
// dropped; each cell keeps value - background.
bool resolve_rect_cells(const std::vector<RectFeature>& rects, cd bg, int n, double pixel,
                        int64_t max_cells, std::vector<RectCell>& out) {
  out.clear();
  const double half = n * pixel / 2.0;
  const double eps = 1e-12 * (2.0 * half);
  struct R { double x0, x1, y0, y1; cd v; };
  std::vector<R> rs;
  for (const auto& f : rects) {
    if (!(f.w > 0 && f.h > 0)) continue;
    const double x0 = std::max(f.x - f.w / 2, -half), x1 = std::min(f.x + f.w / 2, half);
    const double y0 = std::max(f.y - f.h / 2, -half), y1 = std::min(f.y + f.h / 2, half);
    if (x1 - x0 <= 0 || y1 - y0 <= 0) continue;
    rs.push_back({x0, x1, y0, y1, f.value});
  }
  if (rs.empty()) return true;
  std::vector<double> xe{-half, half}, ye{-half, half};
  for (auto& r : rs) {
    xe.push_back(r.x0); xe.push_back(r.x1);
    ye.push_back(r.y0); ye.push_back(r.y1);
  }
  dedup(xe, eps);
  dedup(ye, eps);
  const int64_t nx = (int64_t)xe.size() - 1, ny = (int64_t)ye.size() - 1;
  if (nx < 1 || ny < 1 || nx > max_cells / ny) return false;
  std::vector<cd> cells((size_t)(nx * ny), bg);
  for (auto& r : rs) {
    const int64_t i0 = edge_index(ye, r.y0, eps), i1 = edge_index(ye, r.y1, eps);
    const int64_t j0 = edge_index(xe, r.x0, eps), j1 = edge_index(xe, r.x1, eps);
    for (int64_t i = i0; i < std::min(i1, ny); ++i)
      for (int64_t j = j0; j < std::min(j1, nx); ++j) cells[i * nx + j] = r.v;
  }
  int64_t i = 0;
  while (i < ny) {
    int64_t k = i + 1;
    while (k < ny && std::equal(cells.begin() + k * nx, cells.begin() + (k + 1) * nx,
                                cells.begin() + i * nx))
      ++k;
    int64_t j = 0;
    while (j < nx) {
      const cd v = cells[i * nx + j];
      int64_t e = j + 1;
      while (e < nx && cells[i * nx + e] == v) ++e;
      if (v != bg) out.push_back({xe[j], xe[e], ye[i], ye[k], v - bg});
      j = e;
    }
    i = k;
  }
  return true;
}

}  // namespace huv
