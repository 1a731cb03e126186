// Register-resident FFT building blocks shared by every HIP kernel of the
// imaging path.  Written as host+device templates so the index and twiddle
// logic can be exercised on the CPU; on gfx950 every loop unrolls and every
// twiddle becomes an immediate.
//
// Conventions (crates/highuvlith-core/src/math/fft2d.rs:11-13): forward is
// F[k] = sum_x x[n] e^{-2 pi i k n / N} (unnormalised); the inverse uses
// e^{+2 pi i k n / N} and callers apply any 1/N scale themselves.
#pragma once

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define HUV_HD __host__ __device__ __forceinline__
#else
#define HUV_HD inline
#endif

namespace huvk {

struct alignas(8) c32 {
  float re, im;
};

HUV_HD c32 cmk(float a, float b) { return c32{a, b}; }
HUV_HD c32 cadd(c32 a, c32 b) { return c32{a.re + b.re, a.im + b.im}; }
HUV_HD c32 csub(c32 a, c32 b) { return c32{a.re - b.re, a.im - b.im}; }
HUV_HD c32 cmul(c32 a, c32 b) {
  return c32{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// acc + a * b as four fused multiply-adds (a separate product and sum would
// take six instructions).
HUV_HD c32 cmac(c32 acc, c32 a, c32 b) {
  acc.re = __builtin_fmaf(a.re, b.re, acc.re);
  acc.re = __builtin_fmaf(-a.im, b.im, acc.re);
  acc.im = __builtin_fmaf(a.re, b.im, acc.im);
  acc.im = __builtin_fmaf(a.im, b.re, acc.im);
  return acc;
}
HUV_HD c32 cconjmul(c32 a, c32 b) {  // conj(a) * b
  return c32{a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
HUV_HD c32 cscale(c32 a, float s) { return c32{a.re * s, a.im * s}; }
HUV_HD float cnorm(c32 a) { return a.re * a.re + a.im * a.im; }
// Multiply by +i (inverse) or -i (forward).
template <int SIGN>
HUV_HD c32 cmul_i(c32 a) {
  return SIGN > 0 ? c32{-a.im, a.re} : c32{a.im, -a.re};
}

// ---------------------------------------------------------------------------
// Compile-time twiddles: cos/sin of 2 pi num/den in double precision by
// octant reduction + Taylor series, rounded once to float.
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int k = 1; k < 14; ++k) {
    term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; ++k) {
    term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}
// cos and sin of 2 pi t/den for integer t (exact octant reduction).
constexpr void cos_sin_2pi(long long t, long long den, double& c, double& s) {
  t %= den;
  if (t < 0) t += den;
  // angle = 2 pi t / den = (pi/4) * (8 t / den)
  const long long e = 8 * t;  // octant index = e / den
  const long long oct = e / den;
  const long long rem = e - oct * den;  // angle within octant = (pi/4) rem / den
  const double a = (kPi / 4.0) * (double)rem / (double)den;
  const double ca = taylor_cos(a), sa = taylor_sin(a);
  const double cb = taylor_cos(kPi / 4.0 - a), sb = taylor_sin(kPi / 4.0 - a);
  switch (oct) {
    case 0: c = ca; s = sa; break;        // [0, 45)
    case 1: c = sb; s = cb; break;        // [45, 90): angle = 90 - b
    case 2: c = -sa; s = ca; break;       // [90, 135): angle = 90 + a
    case 3: c = -cb; s = sb; break;       // [135, 180): angle = 180 - b
    case 4: c = -ca; s = -sa; break;      // [180, 225)
    case 5: c = -sb; s = -cb; break;      // [225, 270): angle = 270 - b
    case 6: c = sa; s = -ca; break;       // [270, 315): angle = 270 + a
    default: c = cb; s = -sb; break;      // [315, 360): angle = 360 - b
  }
}
constexpr float twr(long long t, long long den) {
  double c = 0, s = 0;
  cos_sin_2pi(t, den, c, s);
  return (float)c;
}
constexpr float twi(long long t, long long den) {
  double c = 0, s = 0;
  cos_sin_2pi(t, den, c, s);
  return (float)s;
}

// x * e^{SIGN * 2 pi i t / den} with t, den compile-time.
template <int SIGN, long long T, long long DEN>
HUV_HD c32 twiddle(c32 x) {
  constexpr long long tt = ((T % DEN) + DEN) % DEN;
  if constexpr (tt == 0) {
    return x;
  } else if constexpr (4 * tt == DEN) {
    return cmul_i<SIGN>(x);
  } else if constexpr (2 * tt == DEN) {
    return c32{-x.re, -x.im};
  } else if constexpr (4 * tt == 3 * DEN) {
    return cmul_i<-SIGN>(x);
  } else {
    constexpr float c = twr(tt, DEN);
    constexpr float s = SIGN > 0 ? twi(tt, DEN) : -twi(tt, DEN);
    return c32{x.re * c - x.im * s, x.re * s + x.im * c};
  }
}

// ---------------------------------------------------------------------------
// Small DFT butterflies (in place, natural order)
// ---------------------------------------------------------------------------
template <int SIGN>
HUV_HD void dft2(c32& a, c32& b) {
  const c32 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <int SIGN>
HUV_HD void dft4(c32& a0, c32& a1, c32& a2, c32& a3) {
  const c32 s0 = cadd(a0, a2), d0 = csub(a0, a2);
  const c32 s1 = cadd(a1, a3), d1 = cmul_i<SIGN>(csub(a1, a3));
  a0 = cadd(s0, s1);
  a2 = csub(s0, s1);
  a1 = cadd(d0, d1);
  a3 = csub(d0, d1);
}

template <int SIGN>
HUV_HD void dft3(c32& a0, c32& a1, c32& a2) {
  // w = e^{SIGN 2 pi i / 3} = -1/2 + SIGN i sqrt(3)/2
  constexpr float h = 0.86602540378443864676f;
  const c32 s = cadd(a1, a2);
  const c32 d = csub(a1, a2);
  const c32 m = c32{a0.re - 0.5f * s.re, a0.im - 0.5f * s.im};
  // SIGN * i * h * d
  const c32 r = SIGN > 0 ? c32{-h * d.im, h * d.re} : c32{h * d.im, -h * d.re};
  a0 = cadd(a0, s);
  a1 = cadd(m, r);
  a2 = csub(m, r);
}

// ---------------------------------------------------------------------------
// Radix-R DFT on R values held in an array (R in {2, 3, 4, 8, 16}).
// ---------------------------------------------------------------------------
template <int SIGN>
struct DftR {
  template <int R>
  static HUV_HD void run(c32* v) {
    if constexpr (R == 1) {
    } else if constexpr (R == 2) {
      dft2<SIGN>(v[0], v[1]);
    } else if constexpr (R == 3) {
      dft3<SIGN>(v[0], v[1], v[2]);
    } else if constexpr (R == 4) {
      dft4<SIGN>(v[0], v[1], v[2], v[3]);
    } else {
      // R = 2^k >= 8: radix-4 decimation in time, recursively.
      // v[4 j + q] for q in 0..3 are four subsequences of length R/4.
      constexpr int Q = R / 4;
      c32 sub[4][Q];
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int j = 0; j < Q; ++j) sub[q][j] = v[4 * j + q];
#pragma unroll
      for (int q = 0; q < 4; ++q) run<Q>(sub[q]);
      // X[k + Q p] = sum_q w_R^{q k} W_4^{q p} S_q[k]
      apply_tw<Q, 0>(sub);
#pragma unroll
      for (int k = 0; k < Q; ++k) {
        c32 a0 = sub[0][k], a1 = sub[1][k], a2 = sub[2][k], a3 = sub[3][k];
        dft4<SIGN>(a0, a1, a2, a3);
        v[k] = a0;
        v[k + Q] = a1;
        v[k + 2 * Q] = a2;
        v[k + 3 * Q] = a3;
      }
    }
  }
  template <int Q, int K>
  static HUV_HD void apply_tw(c32 (&sub)[4][Q]) {
    if constexpr (K < Q) {
      sub[1][K] = twiddle<SIGN, 1 * K, 4 * Q>(sub[1][K]);
      sub[2][K] = twiddle<SIGN, 2 * K, 4 * Q>(sub[2][K]);
      sub[3][K] = twiddle<SIGN, 3 * K, 4 * Q>(sub[3][K]);
      apply_tw<Q, K + 1>(sub);
    }
  }
};

template <int SIGN, int H, int K>
HUV_HD void tw_half(c32* o) {
  if constexpr (K < H) {
    o[K] = twiddle<SIGN, K, 2 * H>(o[K]);
    tw_half<SIGN, H, K + 1>(o);
  }
}

template <int SIGN, int N>
HUV_HD void fft_reg(c32* v);

template <int SIGN, int L, int K>
HUV_HD void tw_third(c32 (&s)[3][L]) {
  if constexpr (K < L) {
    s[1][K] = twiddle<SIGN, K, 3 * L>(s[1][K]);
    s[2][K] = twiddle<SIGN, 2 * K, 3 * L>(s[2][K]);
    tw_third<SIGN, L, K + 1>(s);
  }
}

// In-register FFT of length N = 2^a or 3 * 2^a (N <= 64), natural order in
// and out.  Lengths with a factor 3 split off a radix-3 decimation-in-time step.
template <int SIGN, int N>
HUV_HD void fft_reg(c32* v) {
  if constexpr (N == 1) {
  } else if constexpr (N % 3 == 0) {
    constexpr int L = N / 3;
    c32 s[3][L];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int j = 0; j < L; ++j) s[q][j] = v[3 * j + q];
#pragma unroll
    for (int q = 0; q < 3; ++q) fft_reg<SIGN, L>(s[q]);
    tw_third<SIGN, L, 0>(s);
#pragma unroll
    for (int k = 0; k < L; ++k) {
      c32 a0 = s[0][k], a1 = s[1][k], a2 = s[2][k];
      dft3<SIGN>(a0, a1, a2);
      v[k] = a0;
      v[k + L] = a1;
      v[k + 2 * L] = a2;
    }
  } else if constexpr (N == 32) {
    // 32 = 2 x 16: decimation in time.
    c32 e[16], o[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      e[j] = v[2 * j];
      o[j] = v[2 * j + 1];
    }
    DftR<SIGN>::template run<16>(e);
    DftR<SIGN>::template run<16>(o);
    tw_half<SIGN, 16, 0>(o);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      v[k] = cadd(e[k], o[k]);
      v[k + 16] = csub(e[k], o[k]);
    }
  } else if constexpr (N == 2) {
    dft2<SIGN>(v[0], v[1]);
  } else {
    DftR<SIGN>::template run<N>(v);
  }
}

}  // namespace huvk
