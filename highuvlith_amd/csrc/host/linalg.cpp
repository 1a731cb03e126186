// Dense complex linear algebra for the SOCS decomposition (FP64, host).
// Follows crates/highuvlith-core/src/math/linalg.rs: tiled Gram product
// (:181-232), Householder tridiagonalisation + diagonal phase normalisation +
// implicit QL with shifts (:324-551, tql2 :343-415), Gram-Schmidt with one
// re-orthogonalisation pass (:575-649) and randomized subspace iteration
// (:657-679).
#include <algorithm>
#include <cmath>
#include <numeric>

#include "engine.hpp"

namespace huv {

namespace {
const int64_t ROW_CHUNK = 256;
const int64_t GRAM_TILE = 48;

// Conjugated dot product sum_i conj(x_i) y_i with two partial accumulators.
cd dotc(const cd* x, const cd* y, int64_t n) {
  double r0 = 0, r1 = 0, i0 = 0, i1 = 0;
  int64_t k = 0;
  for (; k + 1 < n; k += 2) {
    r0 += x[k].real() * y[k].real() + x[k].imag() * y[k].imag();
    i0 += x[k].real() * y[k].imag() - x[k].imag() * y[k].real();
    r1 += x[k + 1].real() * y[k + 1].real() + x[k + 1].imag() * y[k + 1].imag();
    i1 += x[k + 1].real() * y[k + 1].imag() - x[k + 1].imag() * y[k + 1].real();
  }
  if (k < n) {
    r0 += x[k].real() * y[k].real() + x[k].imag() * y[k].imag();
    i0 += x[k].real() * y[k].imag() - x[k].imag() * y[k].real();
  }
  return cd(r0 + r1, i0 + i1);
}

// X * B with B (cols x k) column-major -> rows x k.
ColMatrix mul(const ColMatrix& x, const std::vector<cd>& b, int64_t k) {
  ColMatrix r;
  r.rows = x.rows;
  r.cols = k;
  r.data.assign((size_t)(x.rows * k), cd(0, 0));
  for (int64_t c = 0; c < x.cols; ++c) {
    const cd* xc = x.col(c);
    for (int64_t j = 0; j < k; ++j) {
      const cd bcj = b[(size_t)(j * x.cols + c)];
      if (bcj == cd(0, 0)) continue;
      cd* out = r.col(j);
      for (int64_t i = 0; i < x.rows; ++i) out[i] += bcj * xc[i];
    }
  }
  return r;
}

// X^H * Q -> (x.cols x q.cols).
ColMatrix adjoint_mul(const ColMatrix& x, const ColMatrix& q) {
  ColMatrix r;
  r.rows = x.cols;
  r.cols = q.cols;
  r.data.assign((size_t)(r.rows * r.cols), cd(0, 0));
  for (int64_t c = 0; c < x.cols; ++c)
    for (int64_t j = 0; j < q.cols; ++j) {
      cd acc(0, 0);
      for (int64_t r0 = 0; r0 < x.rows; r0 += ROW_CHUNK) {
        const int64_t r1 = std::min(r0 + ROW_CHUNK, x.rows);
        acc += dotc(x.col(c) + r0, q.col(j) + r0, r1 - r0);
      }
      r.col(j)[c] = acc;
    }
  return r;
}

// splitmix64 + Box-Muller complex Gaussian with E|z|^2 = 1.
struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  cd cgauss() {
    const double u1 = uniform(), u2 = uniform();
    const double r = std::sqrt(-2.0 * std::log(u1));
    return cd(r * std::cos(2 * M_PI * u2), r * std::sin(2 * M_PI * u2)) * M_SQRT1_2;
  }
};

void orthonormalize(ColMatrix& m, Rng& rng) {
  const int64_t rows = m.rows;
  for (int64_t j = 0; j < m.cols; ++j) {
    cd* v = m.col(j);
    const double original = std::sqrt(dotc(v, v, rows).real());
    int attempts = 0;
    while (true) {
      for (int pass = 0; pass < 2 && j > 0; ++pass) {
        std::vector<cd> coeff((size_t)j);
        for (int64_t i = 0; i < j; ++i) coeff[i] = dotc(m.col(i), v, rows);
        for (int64_t i = 0; i < j; ++i) {
          const cd* qi = m.col(i);
          for (int64_t r = 0; r < rows; ++r) v[r] -= coeff[i] * qi[r];
        }
      }
      const double norm = std::sqrt(dotc(v, v, rows).real());
      const double reference = attempts == 0 ? original : 1.0;
      if (norm > 1e-10 * reference && norm > 0.0) {
        const double inv = 1.0 / norm;
        for (int64_t r = 0; r < rows; ++r) v[r] *= inv;
        break;
      }
      ++attempts;
      if (attempts >= 8) break;
      for (int64_t r = 0; r < rows; ++r) v[r] = rng.cgauss();
    }
  }
}

// Implicit QL for a real symmetric tridiagonal (linalg.rs:343-415).
bool tql2(std::vector<double>& d, std::vector<double>& e, std::vector<double>& z) {
  const int64_t n = (int64_t)d.size();
  double dmax = 0, emax = 0;
  for (double v : d) dmax = std::max(dmax, std::fabs(v));
  for (double v : e) emax = std::max(emax, std::fabs(v));
  const double eps = std::numeric_limits<double>::epsilon();
  const double small = eps * (dmax + 2.0 * emax);
  for (int64_t l = 0; l < n; ++l) {
    int iterations = 0;
    while (true) {
      int64_t m = l;
      while (m + 1 < n) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd || std::fabs(e[m]) <= small) break;
        ++m;
      }
      if (m == l) break;
      if (++iterations > 90) return false;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      bool underflow = false;
      int64_t i = m;
      while (i > l) {
        --i;
        const double f = s * e[i];
        const double b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          underflow = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        double* zi = z.data() + i * n;
        double* zi1 = z.data() + (i + 1) * n;
        for (int64_t k = 0; k < n; ++k) {
          const double t = zi1[k];
          zi1[k] = s * zi[k] + c * t;
          zi[k] = c * zi[k] - s * t;
        }
      }
      if (underflow) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  return true;
}

}  // namespace

double ColMatrix::frobenius_norm_sq() const {
  double acc = 0.0;
  const int64_t chunk = ROW_CHUNK * 16;
  for (int64_t k = 0; k < (int64_t)data.size(); k += chunk) {
    const int64_t hi = std::min<int64_t>(k + chunk, (int64_t)data.size());
    double part = 0.0;
    for (int64_t i = k; i < hi; ++i) part += std::norm(data[i]);
    acc += part;
  }
  return acc;
}

std::vector<cd> ColMatrix::gram() const {
  const int64_t n = cols;
  std::vector<cd> h((size_t)(n * n), cd(0, 0));
  const int64_t n_tiles = (n + GRAM_TILE - 1) / GRAM_TILE;
  for (int64_t ta = 0; ta < n_tiles; ++ta)
    for (int64_t tb = ta; tb < n_tiles; ++tb) {
      const int64_t a0 = ta * GRAM_TILE, a1 = std::min((ta + 1) * GRAM_TILE, n);
      const int64_t b0 = tb * GRAM_TILE, b1 = std::min((tb + 1) * GRAM_TILE, n);
      const int64_t wb = b1 - b0;
      std::vector<cd> tile((size_t)((a1 - a0) * wb), cd(0, 0));
      for (int64_t r = 0; r < rows; r += ROW_CHUNK) {
        const int64_t r_end = std::min(r + ROW_CHUNK, rows);
        for (int64_t a = a0; a < a1; ++a) {
          const int64_t bs = ta == tb ? a : b0;
          for (int64_t b = bs; b < b1; ++b)
            tile[(a - a0) * wb + (b - b0)] += dotc(col(a) + r, col(b) + r, r_end - r);
        }
      }
      for (int64_t a = a0; a < a1; ++a) {
        const int64_t bs = ta == tb ? a : b0;
        for (int64_t b = bs; b < b1; ++b) {
          const cd v = tile[(a - a0) * wb + (b - b0)];
          h[b * n + a] = v;            // h[(a, b)] column-major
          h[a * n + b] = std::conj(v);  // h[(b, a)]
        }
      }
    }
  for (int64_t a = 0; a < n; ++a) h[a * n + a] = cd(h[a * n + a].real(), 0.0);
  return h;
}

bool hermitian_eigen_desc(const std::vector<cd>& hin, int64_t n, int64_t count,
                          std::vector<double>& values, std::vector<cd>& vectors) {
  values.clear();
  vectors.clear();
  if (n == 0) return true;
  count = std::min(count, n);
  double scale = 0.0;
  for (const cd& v : hin) scale = std::max(scale, std::abs(v));
  const double inv = scale > 0.0 ? 1.0 / scale : 1.0;
  std::vector<cd> a(hin.size());
  for (size_t i = 0; i < hin.size(); ++i) a[i] = hin[i] * inv;
  auto A = [&](int64_t i, int64_t j) -> cd& { return a[(size_t)(j * n + i)]; };
  std::vector<cd> sub((size_t)std::max<int64_t>(n - 1, 0), cd(0, 0));
  std::vector<std::vector<cd>> reflectors;
  for (int64_t k = 0; k + 2 < n; ++k) {
    const int64_t m = n - k - 1;
    double alpha2 = 0.0;
    for (int64_t i = 0; i < m; ++i) alpha2 += std::norm(A(k + 1 + i, k));
    const double alpha = std::sqrt(alpha2);
    if (alpha == 0.0) {
      sub[k] = cd(0, 0);
      reflectors.emplace_back();
      continue;
    }
    const cd x0 = A(k + 1, k);
    const cd phase = std::abs(x0) > 0.0 ? x0 / std::abs(x0) : cd(1.0, 0.0);
    const cd beta = -phase * alpha;
    std::vector<cd> u((size_t)m);
    for (int64_t i = 0; i < m; ++i) u[i] = A(k + 1 + i, k);
    u[0] -= beta;
    double un = 0.0;
    for (auto& v : u) un += std::norm(v);
    un = std::sqrt(un);
    for (auto& v : u) v /= un;
    std::vector<cd> p((size_t)m, cd(0, 0));
    for (int64_t j = 0; j < m; ++j) {
      const cd uj = u[j] * 2.0;
      for (int64_t i = 0; i < m; ++i) p[i] += A(k + 1 + i, k + 1 + j) * uj;
    }
    const double kk = dotc(u.data(), p.data(), m).real();
    std::vector<cd> q((size_t)m);
    for (int64_t i = 0; i < m; ++i) q[i] = p[i] - u[i] * kk;
    for (int64_t j = 0; j < m; ++j) {
      const cd qj = std::conj(q[j]), uj = std::conj(u[j]);
      for (int64_t i = 0; i < m; ++i) A(k + 1 + i, k + 1 + j) -= u[i] * qj + q[i] * uj;
    }
    sub[k] = beta;
    reflectors.push_back(std::move(u));
  }
  if (n >= 2) sub[n - 2] = A(n - 1, n - 2);
  std::vector<double> d((size_t)n), e((size_t)n, 0.0);
  for (int64_t i = 0; i < n; ++i) d[i] = A(i, i).real();
  for (int64_t i = 0; i + 1 < n; ++i) e[i] = std::abs(sub[i]);
  std::vector<cd> phases((size_t)n, cd(1.0, 0.0));
  for (int64_t k = 0; k + 1 < n; ++k) {
    const double r = std::abs(sub[k]);
    phases[k + 1] = r > 0.0 ? phases[k] * (sub[k] / r) : phases[k];
  }
  std::vector<double> z((size_t)(n * n), 0.0);
  for (int64_t i = 0; i < n; ++i) z[i * n + i] = 1.0;
  if (!tql2(d, e, z)) return false;
  std::vector<int64_t> order((size_t)n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return d[y] < d[x]; });
  values.resize((size_t)n);
  for (int64_t i = 0; i < n; ++i) values[i] = d[order[i]] * scale;
  vectors.assign((size_t)(n * count), cd(0, 0));
  for (int64_t c = 0; c < count; ++c)
    for (int64_t i = 0; i < n; ++i)
      vectors[c * n + i] = phases[i] * z[order[c] * n + i];
  for (int64_t k = (int64_t)reflectors.size() - 1; k >= 0; --k) {
    const auto& u = reflectors[k];
    if (u.empty()) continue;
    const int64_t m = (int64_t)u.size();
    for (int64_t c = 0; c < count; ++c) {
      cd* tail = vectors.data() + c * n + k + 1;
      const cd t = dotc(u.data(), tail, m) * 2.0;
      for (int64_t i = 0; i < m; ++i) tail[i] -= u[i] * t;
    }
  }
  return true;
}

void randomized_eigen(const ColMatrix& x, int64_t l, int power_iters, uint64_t seed,
                      std::vector<double>& values, ColMatrix& u) {
  l = std::max<int64_t>(1, std::min(l, std::min(x.rows, x.cols)));
  Rng rng{seed};
  std::vector<cd> omega((size_t)(x.cols * l));
  for (int64_t c = 0; c < l; ++c)
    for (int64_t r = 0; r < x.cols; ++r) omega[c * x.cols + r] = rng.cgauss();
  ColMatrix q = mul(x, omega, l);
  orthonormalize(q, rng);
  for (int it = 0; it < power_iters; ++it) {
    ColMatrix z = adjoint_mul(x, q);
    orthonormalize(z, rng);
    q = mul(x, z.data, z.cols);
    orthonormalize(q, rng);
  }
  ColMatrix z = adjoint_mul(x, q);
  std::vector<cd> t = z.gram();
  std::vector<cd> w;
  hermitian_eigen_desc(t, l, l, values, w);
  u = mul(q, w, l);
}

}  // namespace huv
