// Pupil functions, pupil fills and adaptive source sampling (FP64, host).
//
// Pupils follow crates/highuvlith-core/src/optics/mod.rs:155-172,310-355 and
// optics/euv.rs:129-165; fills follow source.rs:1178-1248; the sampler
// follows aerial.rs:372-947 step by step (probe + zoom, sign-symmetrised box,
// rotated golden-ratio lattice in the fundamental domain of the box's mirror
// group, 8x8 sub-sample integration of edge cells, graded-cell refinement).
#include <algorithm>
#include <cmath>
#include <limits>

#include "engine.hpp"

namespace huv {

static const double PI = 3.14159265358979323846;

double defocus_phase_in_medium(double z, double na_rho, double nmed, double lambda,
                               bool paraxial) {
  const double s = na_rho / nmed;
  const double s2 = s * s;
  if (paraxial) return PI * z * nmed * s2 / lambda;
  const double cos_t = std::sqrt(std::max(1.0 - s2, 0.0));
  return 2.0 * PI * nmed * z * s2 / ((1.0 + cos_t) * lambda);
}

cd pupil_function(const Optics& o, double fx, double fy, double z, double lambda) {
  const double rho2 = fx * fx + fy * fy;
  if (rho2 > 1.0) return cd(0.0, 0.0);
  const double rho = std::sqrt(rho2);
  if (o.kind == OPTICS_EUV) {
    if (rho < o.central_obscuration) return cd(0.0, 0.0);
    const double phase =
        z == 0.0 ? 0.0 : defocus_phase_in_medium(z, o.na * rho, 1.0, lambda, o.paraxial_defocus);
    return std::polar(std::sqrt(std::max(o.transmission, 0.0)), phase);
  }
  const double phase = z == 0.0 ? 0.0
                                : defocus_phase_in_medium(z, o.na * rho, o.immersion_index,
                                                          lambda, o.paraxial_defocus);
  return std::polar(1.0, phase);
}

// Signed angular distance (deg) in [-180, 180), as the reference's
// ((a - c + 180) rem_euclid 360) - 180.
static double angle_diff_deg(double a, double c) {
  double t = std::fmod(a - c + 180.0, 360.0);
  if (t < 0) t += 360.0;
  return t - 180.0;
}

double illumination_at(const Illumination& s, double x, double y) {
  const double rho = std::sqrt(x * x + y * y);
  switch (s.kind) {
    case ILLUM_CONVENTIONAL:
      return rho <= s.p[0] ? 1.0 : 0.0;
    case ILLUM_ANNULAR:
      return (rho >= s.p[0] && rho <= s.p[1]) ? 1.0 : 0.0;
    case ILLUM_QUADRUPOLE: {
      const double sc = s.p[0], sr = s.p[1], half_open = s.p[2] / 2.0;
      const double angle = std::atan2(y, x) * (180.0 / PI);
      for (double pole : {0.0, 90.0, 180.0, 270.0}) {
        const bool in_pole = std::fabs(angle_diff_deg(angle, pole)) <= half_open;
        const double pr = pole * (PI / 180.0);
        const double cx = sc * std::cos(pr), cy = sc * std::sin(pr);
        const double dist = std::sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy));
        if (in_pole && dist <= sr) return 1.0;
      }
      return 0.0;
    }
    case ILLUM_DIPOLE: {
      const double sc = s.p[0], sr = s.p[1];
      const double orient = s.p[2] * (PI / 180.0);
      for (double pole : {orient, orient + PI}) {
        const double cx = sc * std::cos(pole), cy = sc * std::sin(pole);
        const double dist = std::sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy));
        if (dist <= sr) return 1.0;
      }
      return 0.0;
    }
    case ILLUM_GAUSSIAN: {
      const double sg = s.p[0];
      if (rho <= 1.0 && sg > 0.0) return std::exp(-rho * rho / (2.0 * sg * sg));
      return 0.0;
    }
    case ILLUM_SECTOR_POLES: {
      // Annular sectors sigma_in <= rho <= sigma_out, |angle - pole| <= opening/2,
      // poles at first + k*360/npoles.  Same angular test as the quadrupole fill.
      const double si = s.p[0], so = s.p[1], half_open = s.p[2] / 2.0, first = s.p[3];
      const int np = std::max(1, (int)std::lround(s.p[4]));
      if (!(rho >= si && rho <= so)) return 0.0;
      const double angle = std::atan2(y, x) * (180.0 / PI);
      for (int k = 0; k < np; ++k) {
        const double pole = first + k * (360.0 / np);
        if (std::fabs(angle_diff_deg(angle, pole)) <= half_open) return 1.0;
      }
      return 0.0;
    }
    default:
      return 0.0;
  }
}

// ---------------------------------------------------------------------------
// Adaptive source sampling (aerial.rs:372-947)
// ---------------------------------------------------------------------------
namespace {

const double PROBE_HALF_WIDTH = 2.0;
const int PROBE_POINTS = 129;
const int ZOOM_POINTS = 129;
const int MAX_ZOOM_LEVELS = 6;
const double SUPPORT_THRESHOLD = 1e-4;
const double RESOLVED_CELLS = 24.0;
const double TARGET_STEP_SIGMA = 0.04;
const double MIN_SOURCE_POINTS = 200.0;
const int SUBCELLS = 8;
const int MAX_REFINE = 8;
const int MAX_REFINE_DEPTH = 3;
const double LATTICE_COS = 0.85065080835204;
const double LATTICE_SIN = 0.5257311121191336;
const int MAX_POINTS_PER_AXIS = 128;
const double COHERENT_EXTENT_FRACTION = 1e-3;

struct Probe {
  double bx0, bx1, by0, by1, step_x, step_y;
  int64_t count;
  double cx, cy;
};

double clean(double v) { return (std::isfinite(v) && v > 0.0) ? v : 0.0; }

bool probe_lattice(const Fill& f, double x0, double x1, double y0, double y1, int n,
                   bool centered, Probe& out) {
  const double step_x = (x1 - x0) / (n - 1);
  const double step_y = (y1 - y0) / (n - 1);
  const double half = (n - 1) / 2.0;
  auto coord = [&](double lo, double step, int i) {
    return centered ? (i - half) * step : lo + i * step;
  };
  std::vector<double> values((size_t)n * n);
  for (int64_t k = 0; k < (int64_t)n * n; ++k) {
    const int iy = (int)(k / n), ix = (int)(k % n);
    values[k] = clean(f(coord(x0, step_x, ix), coord(y0, step_y, iy)));
  }
  double imax = 0.0;
  for (double v : values) imax = std::max(imax, v);
  if (imax <= 0.0) return false;
  const double threshold = SUPPORT_THRESHOLD * imax;
  double bx0 = std::numeric_limits<double>::max(), bx1 = std::numeric_limits<double>::lowest();
  double by0 = std::numeric_limits<double>::max(), by1 = std::numeric_limits<double>::lowest();
  int64_t count = 0;
  double wsum = 0, wx = 0, wy = 0;
  for (int64_t k = 0; k < (int64_t)values.size(); ++k) {
    const double v = values[k];
    if (v <= 0.0) continue;
    const double x = coord(x0, step_x, (int)(k % n));
    const double y = coord(y0, step_y, (int)(k / n));
    wsum += v;
    wx += v * x;
    wy += v * y;
    if (v > threshold) {
      ++count;
      bx0 = std::min(bx0, x);
      bx1 = std::max(bx1, x);
      by0 = std::min(by0, y);
      by1 = std::max(by1, y);
    }
  }
  out = Probe{bx0, bx1, by0, by1, step_x, step_y, count, wx / wsum, wy / wsum};
  return true;
}

struct Mirror {
  bool swap, flip_x, flip_y;
  bool operator==(const Mirror& o) const {
    return swap == o.swap && flip_x == o.flip_x && flip_y == o.flip_y;
  }
  void apply(double x, double y, double& ox, double& oy) const {
    double a = swap ? y : x, b = swap ? x : y;
    ox = flip_x ? -a : a;
    oy = flip_y ? -b : b;
  }
  Mirror compose(const Mirror& h) const {
    const bool hx = swap ? h.flip_y : h.flip_x;
    const bool hy = swap ? h.flip_x : h.flip_y;
    return Mirror{(bool)(swap ^ h.swap), (bool)(flip_x ^ hx), (bool)(flip_y ^ hy)};
  }
};

const Mirror M_ID{false, false, false};
const Mirror M_X{false, true, false};
const Mirror M_Y{false, false, true};
const Mirror M_DIAG{true, false, false};

enum WallTest { WALL_XPOS, WALL_YPOS, WALL_BELOW_DIAG };

struct Symmetry {
  bool x, y, diagonal;
  std::vector<std::pair<Mirror, WallTest>> walls() const {
    std::vector<std::pair<Mirror, WallTest>> w;
    if (x && !diagonal) w.push_back({M_X, WALL_XPOS});
    if (y) w.push_back({M_Y, WALL_YPOS});
    if (diagonal) w.push_back({M_DIAG, WALL_BELOW_DIAG});
    return w;
  }
  std::vector<Mirror> group() const {
    std::vector<Mirror> gens;
    for (auto& w : walls()) gens.push_back(w.first);
    std::vector<Mirror> g{M_ID};
    size_t k = 0;
    while (k < g.size()) {
      for (auto& h : gens) {
        Mirror e = g[k].compose(h);
        if (std::find(g.begin(), g.end(), e) == g.end()) g.push_back(e);
      }
      ++k;
    }
    return g;
  }
};

struct CellPoint {
  SourcePoint point;
  bool graded;
  Mirror image;
};

struct CellGeometry {
  const Fill& f;
  double cx, cy, step, hbx, hby;
  std::vector<std::pair<Mirror, WallTest>> walls;

  void to_xy(double u, double v, double& x, double& y) const {
    x = step * (LATTICE_COS * u - LATTICE_SIN * v);
    y = step * (LATTICE_SIN * u + LATTICE_COS * v);
  }
  bool in_domain(double x, double y) const {
    if (!(std::fabs(x) <= hbx && std::fabs(y) <= hby)) return false;
    for (auto& w : walls) {
      switch (w.second) {
        case WALL_XPOS: if (!(x > 0.0)) return false; break;
        case WALL_YPOS: if (!(y > 0.0)) return false; break;
        case WALL_BELOW_DIAG: if (!(y < x)) return false; break;
      }
    }
    return true;
  }
  double eval(const Mirror& g, double x, double y) const {
    double gx, gy;
    g.apply(x, y, gx, gy);
    return clean(f(cx + gx, cy + gy));
  }
  SourcePoint point(const Mirror& g, double x, double y, double w) const {
    double gx, gy;
    g.apply(x, y, gx, gy);
    return SourcePoint{cx + gx, cy + gy, w};
  }

  bool integrate(double u0, double v0, double du, const Mirror& g, CellPoint& out) const {
    const double area = du * du;
    double cxy[2];
    to_xy(u0 + 0.5 * du, v0 + 0.5 * du, cxy[0], cxy[1]);
    const double offs[4][2] = {{0.0, 0.0}, {du, 0.0}, {0.0, du}, {du, du}};
    double corners[4][2];
    bool interior = true;
    for (int c = 0; c < 4; ++c) {
      to_xy(u0 + offs[c][0], v0 + offs[c][1], corners[c][0], corners[c][1]);
      interior = interior && in_domain(corners[c][0], corners[c][1]);
    }
    const double vc = eval(g, cxy[0], cxy[1]);
    if (interior) {
      bool uniform = true;
      for (int c = 0; c < 4 && uniform; ++c)
        uniform = eval(g, corners[c][0], corners[c][1]) == vc;
      if (uniform) {
        if (!(vc > 0.0)) return false;
        out = CellPoint{point(g, cxy[0], cxy[1], vc * area), false, g};
        return true;
      }
    }
    const double sub = (double)SUBCELLS;
    double mass = 0, mx = 0, my = 0;
    int inside = 0, lit = 0;
    double vmin = std::numeric_limits<double>::infinity(), vmax = 0.0;
    for (int k = 0; k < SUBCELLS * SUBCELLS; ++k) {
      const double a = (double)(k % SUBCELLS), b = (double)(k / SUBCELLS);
      double x, y;
      to_xy(u0 + (a + 0.5) / sub * du, v0 + (b + 0.5) / sub * du, x, y);
      if (!in_domain(x, y)) continue;
      ++inside;
      const double v = eval(g, x, y);
      vmin = std::min(vmin, v);
      vmax = std::max(vmax, v);
      if (v > 0.0) {
        mass += v;
        mx += v * x;
        my += v * y;
        ++lit;
      }
    }
    if (mass <= 0.0) return false;
    const bool graded = lit == inside && vmin < vmax;
    SourcePoint p;
    if (interior && lit == SUBCELLS * SUBCELLS && vc > 0.0)
      p = point(g, cxy[0], cxy[1], vc * area);
    else
      p = point(g, mx / mass, my / mass, mass / (sub * sub) * area);
    out = CellPoint{p, graded, g};
    return true;
  }

  void refine(double u0, double v0, double du, const CellPoint& cell, double target, int depth,
              std::vector<SourcePoint>& out) const {
    if (!(cell.graded && cell.point.weight > target) || depth >= MAX_REFINE_DEPTH) {
      out.push_back(cell.point);
      return;
    }
    int r = (int)std::ceil(std::sqrt(cell.point.weight / target));
    r = std::min(std::max(r, 2), MAX_REFINE);
    const double dc = du / r;
    for (int b = 0; b < r; ++b)
      for (int a = 0; a < r; ++a) {
        const double cu = u0 + a * dc, cv = v0 + b * dc;
        CellPoint child;
        if (integrate(cu, cv, dc, cell.image, child))
          refine(cu, cv, dc, child, target, depth + 1, out);
      }
  }
};

std::vector<SourcePoint> lattice_cells(const Fill& f, double cx, double cy, double w, double h,
                                       double step, const Symmetry& sym) {
  CellGeometry geo{f, cx, cy, step, 0.5 * w, 0.5 * h, sym.walls()};
  const double margin = step * std::sqrt(2.0);
  const double reach = std::hypot(w, h) * 0.5 + margin;
  const int64_t n = (int64_t)std::ceil(reach / step) + 1;
  std::vector<std::pair<double, double>> cells;
  for (int64_t j = -n; j < n; ++j)
    for (int64_t i = -n; i < n; ++i) {
      double x, y;
      geo.to_xy(i + 0.5, j + 0.5, x, y);
      const bool near_domain = (!sym.x || sym.diagonal || x > -margin) &&
                               (!sym.y || y > -margin) &&
                               (!sym.diagonal || y < x + margin);
      const bool near_box =
          std::fabs(x) <= 0.5 * w + margin && std::fabs(y) <= 0.5 * h + margin;
      if (near_box && near_domain) cells.push_back({(double)i, (double)j});
    }
  const std::vector<Mirror> group = sym.group();
  const size_t n_img = group.size();
  std::vector<CellPoint> coarse(cells.size() * n_img);
  std::vector<char> present(cells.size() * n_img, 0);
  for (size_t k = 0; k < cells.size() * n_img; ++k)
    present[k] = geo.integrate(cells[k / n_img].first, cells[k / n_img].second, 1.0,
                               group[k % n_img], coarse[k]);
  double total = 0.0;
  for (size_t k = 0; k < coarse.size(); ++k)
    if (present[k]) total += coarse[k].point.weight;
  const double target = total / MIN_SOURCE_POINTS;
  std::vector<SourcePoint> out;
  for (size_t k = 0; k < coarse.size(); ++k)
    if (present[k])
      geo.refine(cells[k / n_img].first, cells[k / n_img].second, 1.0, coarse[k], target, 0,
                 out);
  return out;
}

void normalized_or_centroid(std::vector<SourcePoint>& pts, double cx, double cy) {
  double total = 0.0;
  for (auto& p : pts) total += p.weight;
  if (pts.empty() || std::isnan(total) || total <= 0.0) {
    pts.assign(1, SourcePoint{cx, cy, 1.0});
    return;
  }
  for (auto& p : pts) p.weight /= total;
}

}  // namespace

int sample_intensity(const Fill& f, int points_per_axis, double df_sigma,
                     std::vector<SourcePoint>& out, std::string& err) {
  Probe probe;
  if (!probe_lattice(f, -PROBE_HALF_WIDTH, PROBE_HALF_WIDTH, -PROBE_HALF_WIDTH,
                     PROBE_HALF_WIDTH, PROBE_POINTS, true, probe)) {
    err = "illumination: source pupil fill has no intensity inside |sigma| <= 2";
    return INVALID_PARAMETER;
  }
  for (int lvl = 0; lvl < MAX_ZOOM_LEVELS; ++lvl) {
    const bool rx = probe.bx1 - probe.bx0 >= RESOLVED_CELLS * probe.step_x;
    const bool ry = probe.by1 - probe.by0 >= RESOLVED_CELLS * probe.step_y;
    if (rx && ry) break;
    Probe p;
    if (probe_lattice(f, probe.bx0 - probe.step_x, probe.bx1 + probe.step_x,
                      probe.by0 - probe.step_y, probe.by1 + probe.step_y, ZOOM_POINTS, false, p))
      probe = p;
    else
      break;
  }
  const bool sym_x = std::fabs(probe.bx0 + probe.bx1) <= 0.5 * probe.step_x;
  const bool sym_y = std::fabs(probe.by0 + probe.by1) <= 0.5 * probe.step_y;
  double cx_box, w, cy_box, h;
  if (sym_x) {
    cx_box = 0.0;
    w = 2.0 * (std::max(std::fabs(probe.bx0), std::fabs(probe.bx1)) + probe.step_x);
  } else {
    cx_box = 0.5 * (probe.bx0 + probe.bx1);
    w = probe.bx1 - probe.bx0 + 2.0 * probe.step_x;
  }
  if (sym_y) {
    cy_box = 0.0;
    h = 2.0 * (std::max(std::fabs(probe.by0), std::fabs(probe.by1)) + probe.step_y);
  } else {
    cy_box = 0.5 * (probe.by0 + probe.by1);
    h = probe.by1 - probe.by0 + 2.0 * probe.step_y;
  }
  if (sym_x && sym_y && std::fabs(w - h) <= std::max(probe.step_x, probe.step_y)) {
    const double m = std::max(w, h);
    w = m;
    h = m;
  }
  const double extent = std::max(w, h);
  if (points_per_axis == 1 || extent < COHERENT_EXTENT_FRACTION * df_sigma) {
    out.assign(1, SourcePoint{sym_x ? 0.0 : probe.cx, sym_y ? 0.0 : probe.cy, 1.0});
    return OK;
  }
  if (points_per_axis > 0) {
    const double step = extent / std::max(points_per_axis, 1);
    const int64_t nx = std::max<int64_t>(1, (int64_t)std::llround(w / step));
    const int64_t ny = std::max<int64_t>(1, (int64_t)std::llround(h / step));
    const double hx = w / nx, hy = h / ny;
    std::vector<SourcePoint> pts;
    for (int64_t iy = 0; iy < ny; ++iy) {
      const double y = cy_box + (iy + 0.5 - ny / 2.0) * hy;
      for (int64_t ix = 0; ix < nx; ++ix) {
        const double x = cx_box + (ix + 0.5 - nx / 2.0) * hx;
        const double v = f(x, y);
        if (std::isfinite(v) && v > 0.0) pts.push_back(SourcePoint{x, y, v});
      }
    }
    normalized_or_centroid(pts, probe.cx, probe.cy);
    out = std::move(pts);
    return OK;
  }
  const double area = (double)std::max<int64_t>(probe.count, 1) * probe.step_x * probe.step_y;
  const double step = std::max(std::min(TARGET_STEP_SIGMA, std::sqrt(area / MIN_SOURCE_POINTS)),
                               extent / MAX_POINTS_PER_AXIS);
  Symmetry sym{sym_x, sym_y, sym_x && sym_y && w == h};
  std::vector<SourcePoint> pts = lattice_cells(f, cx_box, cy_box, w, h, step, sym);
  normalized_or_centroid(pts, probe.cx, probe.cy);
  out = std::move(pts);
  return OK;
}

}  // namespace huv
