// HIP kernels of the band-limited SOCS image path (gfx950).
//
// Reference semantics (crates/highuvlith-core/src/aerial.rs:1682-1759):
//   I = (1 - phi) * sum_k lambda_k |IFFT_n(K_k . M)|^2 / I_clear + phi * <.>
// with M = FFT_n(mask) (compute_from_transmittance, aerial.rs:1789-1799).
//
// Every kernel K_k is zero outside the disk |f| <= r, so each field E_k is a
// trigonometric polynomial of per-axis degree B = floor(r) and |E_k|^2 one of
// degree 2B.  Sampling on a coarse m x m grid with m >= 4B + 1 is therefore
// exact (no aliasing): the fields are synthesised and squared on that grid
// and never leave the chip; only the coarse intensity is transformed back
// once per image and evaluated on the n x n output grid.
//
// Stages (one launch each, all batched over clips / images):
//   rows_band    mask rows (real, FP32) -> row spectra, band |kx| <= B kept
//   cols_band    row spectra -> M[ky][kx] band, twisted per x class q
//   socs_fields  K kernels x band -> coarse intensity (fields in registers)
//   coarse_rows  coarse rows -> |gx| <= 2B
//   coarse_cols  -> intensity band I^[gy][gx], 1/m^2, flare on non-DC terms
//   band_image   band -> n x n image rows (C2R), written once to HBM
#include <hip/hip_runtime.h>

#include "fft_lds.hpp"
#include "socs_kernels.hpp"

namespace huvk {

// ---------------------------------------------------------------------------
// rows_band: two real rows packed as z = a + i b, FFT_n forward, keep bins
// k in [0, B] and [n - B, n).  tmp layout: [clip][kk][pair], kk in [0, 2B],
// kk <= B -> k = kk, kk > B -> k = n - (kk - B).  The masks are read once, as
// non-temporal loads, so they do not evict the row spectra that cols_band
// reads next.
// ---------------------------------------------------------------------------
template <int N1, int N2, int TEAMS>
__global__ __launch_bounds__(TEAMS * N2) void k_rows_band(const float* __restrict__ masks, int n,
                                                          int B, const float2* __restrict__ tw_n,
                                                          float2* __restrict__ tmp) {
  constexpr int N = N1 * N2;
  __shared__ c32 scratch[TEAMS][(N1 + 1) * N2];
  const int team = threadIdx.x / N2, t = threadIdx.x % N2;
  const int clip = blockIdx.y;
  const int pair = blockIdx.x * TEAMS + team;
  const int npairs = N / 2;
  const bool active = pair < npairs;
  const float* ra = masks + ((size_t)clip * N + 2 * (size_t)(active ? pair : 0)) * N;
  const float* rb = ra + N;
  const int nk = 2 * B + 1;
  float2* out = tmp + (size_t)clip * nk * npairs;
  fft_four_step<-1, N1, N2>(
      t, scratch[team], tw_n,
      [&](int i) {
        return active ? c32{__builtin_nontemporal_load(ra + i), __builtin_nontemporal_load(rb + i)}
                      : c32{0.f, 0.f};
      },
      [&](int k, c32 v) {
        if (!active) return;
        int kk = -1;
        if (k <= B) kk = k;
        else if (k >= N - B) kk = B + (N - k);
        if (kk >= 0) out[(size_t)kk * npairs + pair] = make_float2(v.re, v.im);
      });
  (void)n;
}

// ---------------------------------------------------------------------------
// cols_band: for column kx in [0, B], assemble the column of row spectra
// from the packed pairs, FFT_n forward along y, keep ky in [-B, B].  Writes
// Mq[clip][kx + BB][ky + BB] = M[ky][kx] * pt and the Hermitian mirror at
// (-kx, -ky); the rest of the (2BB+1)^2 box (BB >= B) is written as zeros.
// pt = pixel transfer sinc(pi kx/n) sinc(pi ky/n) (table sinc_t[k + B], all
// ones when disabled).
// ---------------------------------------------------------------------------
template <int N1, int N2, int TEAMS>
__global__ __launch_bounds__(TEAMS * N2) void k_cols_band(
    const float2* __restrict__ tmp, int B, int BB, const float2* __restrict__ tw_n,
    const float* __restrict__ sinc_t, float2* __restrict__ Mq, float2* __restrict__ Mp) {
  constexpr int N = N1 * N2;
  __shared__ c32 scratch[TEAMS][(N1 + 1) * N2];
  const int team = threadIdx.x / N2, t = threadIdx.x % N2;
  const int clip = blockIdx.y;
  const int kx = blockIdx.x * TEAMS + team;
  const bool active = kx <= BB;
  const bool live = kx <= B;
  const int kxa = live ? kx : 0;
  const int npairs = N / 2;
  const int nk = 2 * B + 1;
  const int W = 2 * BB + 1;
  const float2* col_pos = tmp + ((size_t)clip * nk + kxa) * npairs;
  const float2* col_neg = tmp + ((size_t)clip * nk + (kxa == 0 ? 0 : B + kxa)) * npairs;
  float2* mq = Mq + (size_t)clip * W * W;
  float2* mp = Mp ? Mp + (size_t)clip * (W + 1) * W : nullptr;
  // column-paired copy: (col, row) at [(col + 1) / 2][row][(col + 1) % 2]
  auto put_p = [&](int col, int row, float2 v) {
    mp[((size_t)((col + 1) >> 1) * W + row) * 2 + ((col + 1) & 1)] = v;
  };
  fft_four_step<-1, N1, N2>(
      t, scratch[team], tw_n,
      [&](int i) {
        const int p = i >> 1;
        const float2 zp = col_pos[p], zn = col_neg[p];
        // A = (Z[k] + conj Z[n-k]) / 2 ; Bv = (Z[k] - conj Z[n-k]) / (2i)
        if ((i & 1) == 0) return c32{0.5f * (zp.x + zn.x), 0.5f * (zp.y - zn.y)};
        return c32{0.5f * (zp.y + zn.y), -0.5f * (zp.x - zn.x)};
      },
      [&](int k, c32 v) {
        if (!active) return;
        int ky;
        if (k <= BB) ky = k;
        else if (k >= N - BB) ky = k - N;
        else return;
        if (live && ky >= -B && ky <= B) v = cscale(v, sinc_t[kx + B] * sinc_t[ky + B]);
        else v = c32{0.f, 0.f};
        mq[(size_t)(kx + BB) * W + (ky + BB)] = make_float2(v.re, v.im);
        if (kx > 0) mq[(size_t)(BB - kx) * W + (BB - ky)] = make_float2(v.re, -v.im);
        if (mp) {
          put_p(kx + BB, ky + BB, make_float2(v.re, v.im));
          if (kx > 0) put_p(BB - kx, BB - ky, make_float2(v.re, -v.im));
          if (kx == 0) put_p(-1, ky + BB, make_float2(0.f, 0.f));  // the zero column
        }
      });
}

// ---------------------------------------------------------------------------
// Complex transmittance rasters (compute_from_transmittance with an
// Array2<Complex64> map, aerial.rs:1789-1799): no Hermitian symmetry, so every
// row is its own complex transform and both signs of kx are column passes.
// rows_band_c: row r of clip c (interleaved complex FP32), FFT_n forward,
// keep k in [0, B] and [n - B, n); tmp layout [clip][kk][row], kk as in
// rows_band (kk <= B -> k = kk, kk > B -> k = n - (kk - B)).
// ---------------------------------------------------------------------------
template <int N1, int N2, int TEAMS>
__global__ __launch_bounds__(TEAMS * N2) void k_rows_band_c(const float2* __restrict__ masks,
                                                            int B, const float2* __restrict__ tw_n,
                                                            float2* __restrict__ tmp) {
  constexpr int N = N1 * N2;
  __shared__ c32 scratch[TEAMS][(N1 + 1) * N2];
  const int team = threadIdx.x / N2, t = threadIdx.x % N2;
  const int clip = blockIdx.y;
  const int row = blockIdx.x * TEAMS + team;
  const bool active = row < N;
  const float* r = reinterpret_cast<const float*>(masks + ((size_t)clip * N + (active ? row : 0)) * N);
  const int nk = 2 * B + 1;
  float2* out = tmp + (size_t)clip * nk * N;
  fft_four_step<-1, N1, N2>(
      t, scratch[team], tw_n,
      [&](int i) {
        return active ? c32{__builtin_nontemporal_load(r + 2 * i),
                            __builtin_nontemporal_load(r + 2 * i + 1)}
                      : c32{0.f, 0.f};
      },
      [&](int k, c32 v) {
        if (!active) return;
        int kk = -1;
        if (k <= B) kk = k;
        else if (k >= N - B) kk = B + (N - k);
        if (kk >= 0) out[(size_t)kk * N + row] = make_float2(v.re, v.im);
      });
}

// cols_band_c: box column col = kx + BB for kx in [-BB, BB]: FFT_n forward of
// the row-spectrum column kx along y, keep ky in [-BB, BB]; the box holds
// M[ky][kx] * pt for |kx|, |ky| <= B and zero elsewhere (plain and, when Mp
// is set, column-paired with its zero column in front, as cols_band writes).
template <int N1, int N2, int TEAMS>
__global__ __launch_bounds__(TEAMS * N2) void k_cols_band_c(
    const float2* __restrict__ tmp, int B, int BB, const float2* __restrict__ tw_n,
    const float* __restrict__ sinc_t, float2* __restrict__ Mq, float2* __restrict__ Mp) {
  constexpr int N = N1 * N2;
  __shared__ c32 scratch[TEAMS][(N1 + 1) * N2];
  const int team = threadIdx.x / N2, t = threadIdx.x % N2;
  const int clip = blockIdx.y;
  const int W = 2 * BB + 1;
  const int col = blockIdx.x * TEAMS + team;
  const bool active = col < W;
  const int kx = col - BB;
  const bool live = active && kx >= -B && kx <= B;
  const int kk = !live ? 0 : (kx >= 0 ? kx : B - kx);
  const int nk = 2 * B + 1;
  const float2* src = tmp + ((size_t)clip * nk + kk) * N;
  float2* mq = Mq + (size_t)clip * W * W;
  float2* mp = Mp ? Mp + (size_t)clip * (W + 1) * W : nullptr;
  auto put_p = [&](int c, int r, float2 v) {
    mp[((size_t)((c + 1) >> 1) * W + r) * 2 + ((c + 1) & 1)] = v;
  };
  fft_four_step<-1, N1, N2>(
      t, scratch[team], tw_n,
      [&](int i) {
        if (!live) return c32{0.f, 0.f};
        const float2 z = src[i];
        return c32{z.x, z.y};
      },
      [&](int k, c32 v) {
        if (!active) return;
        int ky;
        if (k <= BB) ky = k;
        else if (k >= N - BB) ky = k - N;
        else return;
        if (live && ky >= -B && ky <= B) v = cscale(v, sinc_t[kx + B] * sinc_t[ky + B]);
        else v = c32{0.f, 0.f};
        mq[(size_t)col * W + (ky + BB)] = make_float2(v.re, v.im);
        if (mp) {
          put_p(col, ky + BB, make_float2(v.re, v.im));
          if (col == 0) put_p(-1, ky + BB, make_float2(0.f, 0.f));  // the zero column
        }
      });
}

// ---------------------------------------------------------------------------
// socs_fields: one work-group per (clip, x class q), all planes of the clip.
// Coarse grid M = 32 Q, Q = R = M / 32 classes and residues, band half-width
// BB = (M - 1) / 4 at compile time (kernel boxes are zero beyond the set's
// own band).  With P = K . M (untwisted band product), a field sample is
//   E[y][x] = sum_{kx, ky} P[kx][ky] e^{2 pi i (kx x + ky y) / M}.
// x group, band row iy (ky = iy - BB), for the columns x = Q j + q:
//   c[h] = sum_t P_q[h + 32 t][ky],  P_q[kx] = K[kx] e^{2 pi i kx q / M} . M[kx]
//   T[iy][j] = IFFT_32(c)[j]
// (the class twist e^{2 pi i kx q / M} is carried by the kernel stack, one
// pre-twisted copy per class, so the x pass is a plain multiply-accumulate)
// y group, column j and residue r, for the rows y = R j' + r:
//   d[h] = e^{2 pi i h r / M} sum_t T[h + 32 t][j] e^{2 pi i t r / R}
//   E[R j' + r][Q j + q] = IFFT_32(d)[j'];  acc += lambda |E|^2
// The groups run software-pipelined over a double-buffered LDS tile (step s
// synthesises kernel s while kernel s - 1 is squared and summed), one barrier
// per step.  Coarse intensity layout: class-major rows, coarse[img][y][q][j]
// for x = Q j + q, so the 32 lanes of a residue store one contiguous 128-byte
// run per row (coarse_rows reads the row back in x order).
// Work-groups are ordered so that the Q classes of one clip share an XCD
// (consecutive work-group ids are dealt round-robin over the 8 XCDs).
// ---------------------------------------------------------------------------
// floor(a / 32)
constexpr int fdiv32(int a) { return a >= 0 ? a / 32 : -((-a + 31) / 32); }

// XS = 1: one x lane per band row, all columns.  XS = 2: two lanes per row
// (lanes l and l + 32 of one wave); lane p takes the columns whose bin
// h = kx mod 32 lies in [16 p, 16 p + 16), in NR runs of 16,
//   kx = 16 p + 32 (xt0(p) + r) + i,  r < NR, i < 16,
// and the pair finishes the 32-point transform with one radix-2 step.  The
// only column outside the band is kx = -BB - 1 (col -1, the first column of
// one of the lanes).  This form reads column-paired band spectra and kernel
// stacks, [(col + 1) / 2][row][(col + 1) % 2] with col in [-1, W) (col -1
// zero), one 16-byte load per lane for two columns.
template <int M, int XS = 1>
struct CoarseCfg {
  static constexpr int F = 32;  // transform length of both passes
  static constexpr int Q = M / F;
  static constexpr int R = Q;
  static constexpr int BB = (M - 1) / 4;
  static constexpr int W = 2 * BB + 1;
  static constexpr int TP = F + 1;
  static constexpr int TMIN = -((BB + F - 1) / F);  // fold range of k = h + 32 t
  static constexpr int TMAX = BB / F;
  static constexpr int YT = F * R;
  static constexpr int XOFF = ((YT + 63) / 64) * 64;
  static constexpr int XT = XS == 1 ? ((W + 63) / 64) * 64 : ((W + 31) / 32) * 64;
  static constexpr int THREADS = XOFF + XT;
  static constexpr int xt0(int p) { return -((BB + 16 * p + 15) >> 5); }
  static constexpr int xcol0(int p) { return 16 * p + 32 * xt0(p) + BB; }
  static constexpr int nr() {
    const int a = fdiv32(BB) - xt0(0) + 1, b = fdiv32(BB - 16) - xt0(1) + 1;
    return a > b ? a : b;
  }
  static constexpr int NR = nr();
  static constexpr bool SPLIT_OK =
      xcol0(0) >= -1 && xcol0(1) >= -1 && xcol0(0) + 32 * (NR - 1) + 15 <= W - 1 &&
      xcol0(1) + 32 * (NR - 1) + 15 <= W - 1;
  static constexpr int LDS_C32 = 2 * W * TP + R * F + F;
};
static_assert(CoarseCfg<64, 2>::SPLIT_OK && CoarseCfg<128, 2>::SPLIT_OK &&
                  CoarseCfg<192, 2>::SPLIT_OK && CoarseCfg<256, 2>::SPLIT_OK,
              "two-lane x pass: columns must lie in [-1, W)");

constexpr int kXcds = 8;

// e^{+2 pi i e / M} from the forward W_M table.
__device__ __forceinline__ c32 tw_pos(const float2* __restrict__ tw_m, int e, int M) {
  e %= M;
  if (e < 0) e += M;
  return tw_lookup<+1>(tw_m, e);
}

// Raw buffer resource over `count` complex values (gfx9 word 3; reads past
// the end return zero).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t box_rsrc(const float2* p, int count) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(p), (short)0,
                                           count * (int)sizeof(float2), 0x00020000);
}

__device__ __forceinline__ c32 load_c32(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
  const float2 f = __builtin_bit_cast(float2, v);
  return c32{f.x, f.y};
}

struct c32x2 {
  c32 a, b;
};

__device__ __forceinline__ c32x2 load_c32x2(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
  const float4 f = __builtin_bit_cast(float4, v);
  return c32x2{c32{f.x, f.y}, c32{f.z, f.w}};
}

// XNB: band columns per load block of the x pass (a tuning parameter).
// kern: [Q][nplanes * Kmax][W][W], class q's copy twisted by e^{2 pi i kx q / M}.
// S > 1 (small batches): work-group split j takes the kernels [j ks, j ks + ks)
// of every plane (ks = ceil(Kmax / S), every split non-empty) and writes its
// partial coarse images at coarse + j * nclips * nplanes * M^2; coarse_rows
// sums the S partials in split order.
template <int M, int XNB, int XS>
__global__ __launch_bounds__((CoarseCfg<M, XS>::THREADS)) __attribute__((amdgpu_waves_per_eu(4))) void k_socs_fields(
    const float2* __restrict__ Mq, const float2* __restrict__ kern, const float* __restrict__ lam,
    int nclips, int nplanes, int Kmax, int S, const float2* __restrict__ tw_m,
    float* __restrict__ coarse) {
  using C = CoarseCfg<M, XS>;
  constexpr int F = C::F, Q = C::Q, R = C::R, BB = C::BB, W = C::W, TP = C::TP;
  constexpr int TMIN = C::TMIN, TMAX = C::TMAX, NT = TMAX - TMIN + 1;
  extern __shared__ c32 lds[];
  c32* T = lds;                    // 2 x W x TP
  c32* twy = lds + 2 * W * TP;     // [r][h] = e^{2 pi i h r / M}
  c32* tw32 = twy + R * F;         // [k] = e^{2 pi i k / 32}
  const int L = blockIdx.x;
  const int slot = L / kXcds;
  const int q = slot % Q;
  const int sj = (slot / Q) % S;
  const int clip = (slot / (Q * S)) * kXcds + L % kXcds;
  if (clip >= nclips) return;  // whole work-group, before any barrier
  const int tid = threadIdx.x;
  for (int i = tid; i < R * F; i += C::THREADS) twy[i] = tw_pos(tw_m, (i % F) * (i / F), M);
  for (int i = tid; i < F; i += C::THREADS) tw32[i] = tw_pos(tw_m, i * Q, M);
  // band spectrum box of this clip (column-paired for XS = 2)
  constexpr int BOX = XS == 1 ? W * W : (W + 1) * W;
  const __amdgpu_buffer_rsrc_t mr = box_rsrc(Mq + (size_t)clip * BOX, BOX);
  const int steps = nplanes * Kmax;
  // this split's kernels k0 .. k0 + kc - 1 of each plane; local step l is
  // plane l / kc, kernel k0 + l % kc
  const int ks = (Kmax + S - 1) / S, k0 = sj * ks, kc = min(ks, Kmax - k0);
  const int lsteps = nplanes * kc;
  auto gstep = [&](int l) { return (l / kc) * Kmax + k0 + l % kc; };
  const bool is_y = tid < C::YT;
  const int j = tid % F, r = tid / F;
  const int ix = tid - C::XOFF;  // x lane
  const int part = XS == 1 ? 0 : (ix >> 5) & 1;
  const int iy = XS == 1 ? ix : (ix < 0 ? -1 : 32 * (ix >> 6) + (ix & 31));
  // y fold factors e^{2 pi i t r / R} (M / R = 32)
  c32 wf[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) wf[t] = tw_pos(tw_m, (t + TMIN) * (is_y ? r : 0) * F, M);
  __syncthreads();
  // The two groups run separate loops (wave-uniform split) with the same
  // number of barriers, so their register live ranges do not overlap.
  if (ix >= 0) {
    for (int s = 0; s <= lsteps; ++s) {
      if (s < lsteps && iy < W) {
        const int sg = gstep(s);
        c32* Tr = T + (s & 1) * W * TP + iy * TP;
        if constexpr (XS == 1) {
          // Buffer loads: one VGPR offset per lane, the column offset in an SGPR.
          const int vo = iy * (int)sizeof(float2);
          const __amdgpu_buffer_rsrc_t kr =
              box_rsrc(kern + ((size_t)q * steps + sg) * W * W, W * W);
          c32 c[F];
#pragma unroll
          for (int h = 0; h < F; ++h) c[h] = c32{0.f, 0.f};
          // Columns kx = -BB..BB in blocks of XNB; the loads of block b + 1
          // are issued before block b is consumed.
          constexpr int NBLK = (W + XNB - 1) / XNB;
          c32 ka[2][XNB], ma[2][XNB];
#pragma unroll
          for (int i = 0; i < XNB; ++i) {
            const int so = i * W * (int)sizeof(float2);
            ka[0][i] = load_c32(kr, vo, so);
            ma[0][i] = load_c32(mr, vo, so);
          }
#pragma unroll
          for (int blk = 0; blk < NBLK; ++blk) {
            __builtin_amdgcn_sched_barrier(0);
            if (blk + 1 < NBLK) {
#pragma unroll
              for (int i = 0; i < XNB; ++i) {
                const int col = (blk + 1) * XNB + i;
                if (col < W) {
                  const int so = col * W * (int)sizeof(float2);
                  ka[(blk + 1) & 1][i] = load_c32(kr, vo, so);
                  ma[(blk + 1) & 1][i] = load_c32(mr, vo, so);
                }
              }
            }
#pragma unroll
            for (int i = 0; i < XNB; ++i) {
              const int col = blk * XNB + i;
              if (col < W) {
                const int kx = col - BB;
                const int h = ((kx % F) + F) % F;
                c[h] = cmac(c[h], ka[blk & 1][i], ma[blk & 1][i]);
              }
            }
          }
          fft_reg<+1, F>(c);
#pragma unroll
          for (int jj = 0; jj < F; ++jj) Tr[jj] = c[jj];
        } else {
          // Lane `part` streams its NR runs of 16 columns from col0, two
          // columns (one pair) per 16-byte load.
          constexpr int NP = 8 * C::NR, XP = XNB / 2;
          const int col0 = part ? C::xcol0(1) : C::xcol0(0);
          const int vo = (((col0 + 1) >> 1) * W + iy) * (int)(2 * sizeof(float2));
          const __amdgpu_buffer_rsrc_t kr =
              box_rsrc(kern + ((size_t)q * steps + sg) * BOX, BOX);
          c32 c[F / 2];  // bins 16 part + h
#pragma unroll
          for (int h = 0; h < F / 2; ++h) c[h] = c32{0.f, 0.f};
          constexpr int NBLK = (NP + XP - 1) / XP;
          c32x2 ka[2][XP], ma[2][XP];
#pragma unroll
          for (int i = 0; i < XP; ++i) {
            const int so = ((i / 8) * 16 + i % 8) * W * (int)(2 * sizeof(float2));
            ka[0][i] = load_c32x2(kr, vo, so);
            ma[0][i] = load_c32x2(mr, vo, so);
          }
#pragma unroll
          for (int blk = 0; blk < NBLK; ++blk) {
            __builtin_amdgcn_sched_barrier(0);
            if (blk + 1 < NBLK) {
#pragma unroll
              for (int i = 0; i < XP; ++i) {
                const int k = (blk + 1) * XP + i;
                if (k < NP) {
                  const int so = ((k / 8) * 16 + k % 8) * W * (int)(2 * sizeof(float2));
                  ka[(blk + 1) & 1][i] = load_c32x2(kr, vo, so);
                  ma[(blk + 1) & 1][i] = load_c32x2(mr, vo, so);
                }
              }
            }
#pragma unroll
            for (int i = 0; i < XP; ++i) {
              const int k = blk * XP + i;
              if (k < NP) {
                const int h = 2 * (k % 8);
                c[h] = cmac(c[h], ka[blk & 1][i].a, ma[blk & 1][i].a);
                c[h + 1] = cmac(c[h + 1], ka[blk & 1][i].b, ma[blk & 1][i].b);
              }
            }
          }
          // radix-2 step across the lane pair through the tile row:
          //   lane 0: FFT_16(c[h] + c[h + 16])                        -> outputs 2k
          //   lane 1: FFT_16((c[h] - c[h + 16]) e^{2 pi i h / 32})    -> outputs 2k + 1
#pragma unroll
          for (int h = 0; h < F / 2; ++h) Tr[F / 2 * part + h] = c[h];
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
          const float sg = part ? -1.f : 1.f;
#pragma unroll
          for (int h = 0; h < F / 2; ++h) {
            const c32 o = Tr[F / 2 * (1 - part) + h];
            c[h] = c32{__builtin_fmaf(sg, c[h].re, o.re), __builtin_fmaf(sg, c[h].im, o.im)};
          }
#pragma unroll
          for (int h = 1; h < F / 2; ++h) c[h] = cmul(c[h], tw32[h * part]);
          fft_reg<+1, F / 2>(c);
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
#pragma unroll
          for (int k = 0; k < F / 2; ++k) Tr[2 * k + part] = c[k];
        }
      }
      __syncthreads();
    }
  } else {
    float acc[F];
#pragma unroll
    for (int i = 0; i < F; ++i) acc[i] = 0.f;
    for (int s = 0; s <= lsteps; ++s) {
      if (is_y && s >= 1) {
        const int sp = s - 1;
        // Re-read the twiddles every step rather than holding them in registers.
        int tyo = r * F;
        asm volatile("" : "+v"(tyo));
        const c32* ty = twy + tyo;
        const c32* Tb = T + (sp & 1) * W * TP;
        c32 d[F];
#pragma unroll
        for (int h = 0; h < F; ++h) {
          c32 v{0.f, 0.f};
#pragma unroll
          for (int t = TMIN; t <= TMAX; ++t) {
            const int ky = h + F * t;
            if (ky >= -BB && ky <= BB) {
              const c32 e = Tb[(ky + BB) * TP + j];
              v = t == 0 ? cadd(v, e) : cmac(v, e, wf[t - TMIN]);
            }
          }
          d[h] = cmul(v, ty[h]);
        }
        fft_reg<+1, F>(d);
        const float l = lam[gstep(sp)];
#pragma unroll
        for (int i = 0; i < F; ++i) acc[i] += l * cnorm(d[i]);
        if (sp % kc == kc - 1) {
          // Vector buffer stores: one VGPR offset, the row offset in an SGPR.
          const __amdgpu_buffer_rsrc_t orc = __builtin_amdgcn_make_buffer_rsrc(
              coarse + (((size_t)sj * nclips + clip) * nplanes + sp / kc) * M * M, (short)0,
              M * M * (int)sizeof(float), 0x00020000);
          const int vo = (r * M + F * q + j) * (int)sizeof(float);
#pragma unroll
          for (int i = 0; i < F; ++i) {
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, acc[i]), orc, vo,
                                                  R * i * M * (int)sizeof(float), 0);
            acc[i] = 0.f;
          }
        }
      }
      __syncthreads();
    }
  }
}

// Four-step factorisation M = N1 x N2 of the coarse transforms.
template <int M>
struct CoarseFft {
  static constexpr int N1 = M == 32 ? 4 : M == 64 ? 8 : M == 128 ? 8 : M == 192 ? 12 : 16;
  static constexpr int N2 = M / N1;
};

// ---------------------------------------------------------------------------
// coarse_rows: coarse row pairs -> FFT_M forward, keep gx in [0, G] and
// [M - G, M) (G = 2B).  tmp1 layout [img][kk][pair] with kk in [0, 2G].
// ---------------------------------------------------------------------------
// coarse[i] = sum_j coarse[j * count + i], j = 0 .. S-1 in order (float4 lanes).
__global__ void k_sum_splits(float* __restrict__ coarse, int S, int64_t count4) {
  float4* c4 = reinterpret_cast<float4*>(coarse);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count4;
       i += (int64_t)gridDim.x * blockDim.x) {
    float4 v = c4[i];
    for (int j = 1; j < S; ++j) {
      const float4 w = c4[j * count4 + i];
      v.x += w.x;
      v.y += w.y;
      v.z += w.z;
      v.w += w.w;
    }
    c4[i] = v;
  }
}

template <int M, int TEAMS>
__global__ __launch_bounds__(TEAMS * CoarseFft<M>::N2) void k_coarse_rows(
    const float* __restrict__ coarse, int B, const float2* __restrict__ tw_m,
    float2* __restrict__ tmp1) {
  constexpr int N1 = CoarseFft<M>::N1, N2 = CoarseFft<M>::N2, Q = M / 32;
  __shared__ c32 scratch[TEAMS][(N1 + 1) * N2];
  const int team = threadIdx.x / N2, t = threadIdx.x % N2;
  const int img = blockIdx.y;
  const int pair = blockIdx.x * TEAMS + team;
  const int npairs = M / 2;
  const bool active = pair < npairs;
  const float* ra = coarse + (size_t)img * M * M + (size_t)2 * (active ? pair : 0) * M;
  const float* rb = ra + M;
  const int G = 2 * B;
  const int nk = 2 * G + 1;
  float2* out = tmp1 + (size_t)img * nk * npairs;
  fft_four_step<-1, N1, N2>(
      t, scratch[team], tw_m,
      [&](int x) {
        // class-major rows: x = Q j + q is stored at 32 q + j (socs_fields)
        const int at = (x % Q) * 32 + x / Q;
        return active ? c32{ra[at], rb[at]} : c32{0.f, 0.f};
      },
      [&](int k, c32 v) {
        if (!active) return;
        int kk = -1;
        if (k <= G) kk = k;
        else if (k >= M - G) kk = G + (M - k);
        if (kk >= 0) out[(size_t)kk * npairs + pair] = make_float2(v.re, v.im);
      });
}

// coarse_cols: column gx in [0, G]: separate the packed pairs, FFT_M along y,
// keep gy in [-G, G]; Ib[img][gy + G][gx] = v / M^2, times (1 - phi) except DC.
template <int M, int TEAMS>
__global__ __launch_bounds__(TEAMS * CoarseFft<M>::N2) void k_coarse_cols(
    const float2* __restrict__ tmp1, int B, float flare, const float2* __restrict__ tw_m,
    float2* __restrict__ Ib) {
  constexpr int N1 = CoarseFft<M>::N1, N2 = CoarseFft<M>::N2;
  __shared__ c32 scratch[TEAMS][(N1 + 1) * N2];
  const int team = threadIdx.x / N2, t = threadIdx.x % N2;
  const int img = blockIdx.y;
  const int G = 2 * B;
  const int gx = blockIdx.x * TEAMS + team;
  const bool active = gx <= G;
  const int gxa = active ? gx : 0;
  const int npairs = M / 2;
  const int nk = 2 * G + 1;
  const float2* col_pos = tmp1 + ((size_t)img * nk + gxa) * npairs;
  const float2* col_neg = tmp1 + ((size_t)img * nk + (gxa == 0 ? 0 : G + gxa)) * npairs;
  float2* out = Ib + (size_t)img * nk * (G + 1) + gxa;
  const float inv = 1.0f / ((float)M * (float)M);
  fft_four_step<-1, N1, N2>(
      t, scratch[team], tw_m,
      [&](int i) {
        const int p = i >> 1;
        const float2 zp = col_pos[p], zn = col_neg[p];
        if ((i & 1) == 0) return c32{0.5f * (zp.x + zn.x), 0.5f * (zp.y - zn.y)};
        return c32{0.5f * (zp.y + zn.y), -0.5f * (zp.x - zn.x)};
      },
      [&](int k, c32 v) {
        if (!active) return;
        int gy;
        if (k <= G) gy = k;
        else if (k >= M - G) gy = k - M;
        else return;
        const float s = (gx == 0 && gy == 0) ? inv : inv * (1.0f - flare);
        out[(size_t)(gy + G) * (G + 1)] = make_float2(v.re * s, v.im * s);
      });
}

// ---------------------------------------------------------------------------
// coarse_fused: coarse_rows and coarse_cols of one image in one work-group,
// the row spectra kept in LDS ((M/2) x (M + 1) complex; M = 64, 128, 192).
// Transforms M = N1 x N2 run in place in the LDS rows, teams of N2 threads:
//  row pass, pair y2 (z = row 2 y2 + i row 2 y2 + 1, x = Q j + q read from
//    the class-major image): thread n1 < N1 takes x = n1 + N1 n2, FFT_N2,
//    twiddle W_M^{n1 k2}, to slot n1 + N1 k2; after a barrier thread k2
//    takes slots n1 + N1 k2 (n1 < N1), FFT_N1, back to the same slots: the
//    pair's frequency k = k2 + N2 k1 sits at slot k1 + N1 k2.
//  column pass, column gx in [0, G] (G = 2B): c[2 y2] = A, c[2 y2 + 1] = B
//    separated from Z[gx] and Z[M - gx] of pair y2, FFT_M over y in the same
//    steps with the column's two slots per pair as the exchange area (slots
//    of frequency M/2, never kept since G < M/2, for gx = 0); gy in [-G, G]
//    kept: Ib[img][gy + G][gx] = v / M^2, times (1 - flare) except at the
//    origin.  Every read of a column's slots precedes the barrier after
//    which they are rewritten; columns never share slots.
// ---------------------------------------------------------------------------
template <int M, int THREADS>
__global__ __launch_bounds__(THREADS) void k_coarse_fused(const float* __restrict__ coarse, int B,
                                                          float flare,
                                                          const float2* __restrict__ tw_m,
                                                          float2* __restrict__ Ib) {
  constexpr int N1 = CoarseFft<M>::N1, N2 = CoarseFft<M>::N2, Q = M / 32;
  constexpr int NP = M / 2, PS = M + 1, TEAMS = THREADS / N2;
  extern __shared__ c32 zs[];  // [NP][PS]
  const int img = blockIdx.x;
  const int team = threadIdx.x / N2, t = threadIdx.x % N2;
  const int G = 2 * B;
  auto slot = [](int k) { return k / N2 + N1 * (k % N2); };
  const float* base = coarse + (size_t)img * M * M;
  for (int r0 = 0; r0 < NP; r0 += TEAMS) {
    const int y2 = r0 + team;
    const bool act = y2 < NP;
    c32* row = zs + (act ? y2 : 0) * PS;
    if (act && t < N1) {
      const float* ra = base + (size_t)(2 * y2) * M;
      const float* rb = ra + M;
      c32 v[N2];
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) {
        const int x = t + N1 * n2;
        const int at = (x % Q) * 32 + x / Q;  // class-major position of x
        v[n2] = c32{ra[at], rb[at]};
      }
      fft_small<-1, N2>(v);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        if (k2 > 0) v[k2] = cmul(v[k2], tw_lookup<-1>(tw_m, (t * k2) % M));
        row[t + N1 * k2] = v[k2];
      }
    }
    __syncthreads();
    if (act) {
      c32 u[N1];
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) u[n1] = row[n1 + N1 * t];
      fft_small<-1, N1>(u);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) row[k1 + N1 * t] = u[k1];
    }
    __syncthreads();
  }
  const float inv = 1.0f / ((float)M * (float)M);
  const int nk = 2 * G + 1;
  float2* out = Ib + (size_t)img * nk * (G + 1);
  for (int c0 = 0; c0 <= G; c0 += TEAMS) {
    const int gx = c0 + team;
    const bool act = gx <= G;
    const int gxa = act ? gx : 0;
    const int sp = slot(gxa), sm = slot((M - gxa) % M), sw = gxa == 0 ? slot(M / 2) : sm;
    // exchange-area slot of column value y
    auto cs = [&](int y) -> c32* { return zs + (y >> 1) * PS + ((y & 1) ? sw : sp); };
    c32 v[N2];
    if (act && t < N1) {
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) {
        const int y = t + N1 * n2;
        const c32 zp = zs[(y >> 1) * PS + sp], zn = zs[(y >> 1) * PS + sm];
        // A = (Z[k] + conj Z[M-k]) / 2 ; B = (Z[k] - conj Z[M-k]) / (2i)
        v[n2] = (y & 1) == 0 ? c32{0.5f * (zp.re + zn.re), 0.5f * (zp.im - zn.im)}
                             : c32{0.5f * (zp.im + zn.im), -0.5f * (zp.re - zn.re)};
      }
      fft_small<-1, N2>(v);
#pragma unroll
      for (int k2 = 1; k2 < N2; ++k2) v[k2] = cmul(v[k2], tw_lookup<-1>(tw_m, (t * k2) % M));
    }
    __syncthreads();
    if (act && t < N1) {
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) *cs(t + N1 * k2) = v[k2];
    }
    __syncthreads();
    if (act) {
      c32 u[N1];
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) u[n1] = *cs(n1 + N1 * t);
      fft_small<-1, N1>(u);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) {
        const int k = t + N2 * k1;
        int gy;
        if (k <= G) gy = k;
        else if (k >= M - G) gy = k - M;
        else continue;
        const float s = (gx == 0 && gy == 0) ? inv : inv * (1.0f - flare);
        out[(size_t)(gy + G) * (G + 1) + gx] = make_float2(u[k1].re * s, u[k1].im * s);
      }
    }
  }
}

// v_permlane32_swap on two floats: afterwards the upper 32 lanes of a hold
// the lower lanes of b and the lower 32 lanes of b the upper lanes of a (the
// other halves keep their values).  The results are taken out of the
// returned vector as integers before the bit cast: a bit cast applied to a
// vector element directly reads element 0.
__device__ __forceinline__ void swap32(float& a, float& b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(int, a),
                                                  __builtin_bit_cast(int, b), false, false);
  const int r0 = r[0], r1 = r[1];
  a = __builtin_bit_cast(float, r0);
  b = __builtin_bit_cast(float, r1);
}

// u[k] *= e^{SIGN 2 pi i k / (2 L)}, k < L, in the lanes where `on` holds.
template <int SIGN, int L, int K = 0>
__device__ __forceinline__ void tw_half_if(c32* u, bool on) {
  if constexpr (K < L) {
    const c32 y = twiddle<SIGN, K, 2 * L>(u[K]);
    u[K] = on ? y : u[K];
    tw_half_if<SIGN, L, K + 1>(u, on);
  }
}

// ---------------------------------------------------------------------------
// band_image: the n x n real image of a Hermitian half band Ib[gy + G][gx]
// (gx in [0, G]), I[y][x] = sum_{g} Ib[g] e^{2 pi i (gx x + gy y) / n}.
// One work-group per (image, row class p) computes the 32 rows y = P j + p
// (P = n / 32):
//   phase 1, thread gx:  U[j][gx] = IFFT_32(u)[j],
//       u[h] = e^{2 pi i h p / n} sum_t Ib[h + 32 t][gx] e^{2 pi i t p / P}
//   phase 2, task (row pair, column class c), C = n / 32 classes, z = a + i b:
//       X[C i + c] = IFFT_32(d)[i],
//       d[h] = e^{2 pi i h c / n} sum_t z[h + 32 t] e^{2 pi i t c / C}
//   with z[g] = Ua[g] + i Ub[g] (g >= 0), conj Ua[-g] + i conj Ub[-g] (g < 0);
//   row a takes Re X, row b Im X.  Writes are coalesced across classes.  The
//   classes c and c + C/2 share their fold (even and odd t, two lanes per
//   pair of classes, one permlane32 swap to regroup by class).
// Folding ranges t in [TMIN, TMAX] cover |g| <= G with 32 (TMAX) <= G <=
// 32 (TMAX + 1) - 1 (TMIN = -TMAX - 1), so only the outermost t of either
// sign can fall outside the band; U rows are zero-padded to |g| <= 32 |TMIN|
// so phase 2 needs no band test at all.
// Work-groups of one image are dealt to one XCD so the band stays in its L2.
// ---------------------------------------------------------------------------
// Fidelity of one pixel (ilt.rs:564-605): the residual of the dosed image
// against the target, sigmoid-resist (KIND 1) or plain (0); returns r^2 and
// sets g = dC/dI (unweighted by the condition weight).
template <int KIND>
__device__ __forceinline__ double fidelity_term(const FidelityEpi& e, float dose, float v, float z,
                                                float& g) {
  const float di = dose * v;
  float r;
  if constexpr (KIND == 0) {
    r = di - z;
    g = 2.0f * r * dose * e.inv_n;
  } else {
    const float rr = 1.0f / (1.0f + expf(-e.steepness * (di - e.threshold)));
    r = rr - z;
    g = 2.0f * r * e.steepness * rr * (1.0f - rr) * dose * e.inv_n;
  }
  return (double)r * (double)r;
}

// The fidelity of one band_image task's 64 pixels (rows a and b at byte
// offsets va, vb, column offsets C i): target values read 8 columns at a
// time, G written when WITH_G; returns the task's sum of r^2.
template <int KIND, bool WITH_G>
__device__ __forceinline__ double fidelity_task(const FidelityEpi& e, float dose, const c32* d,
                                                int C, int va, int vb,
                                                __amdgpu_buffer_rsrc_t zrc,
                                                __amdgpu_buffer_rsrc_t grc) {
  double acc = 0.0;
#pragma unroll
  for (int i0 = 0; i0 < 32; i0 += 8) {
    __builtin_amdgcn_sched_barrier(0);
    float za[8], zb[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int so = C * (i0 + k) * (int)sizeof(float);
      za[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(zrc, va, so, 0));
      zb[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(zrc, vb, so, 0));
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      float ga, gb;
      acc += fidelity_term<KIND>(e, dose, d[i0 + k].re, za[k], ga);
      acc += fidelity_term<KIND>(e, dose, d[i0 + k].im, zb[k], gb);
      if constexpr (WITH_G) {
        const int so = C * (i0 + k) * (int)sizeof(float);
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, ga), grc, va, so, 0);
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, gb), grc, vb, so, 0);
      }
    }
  }
  return acc;
}

// The ILT chain rule on one band_image task's 64 pixels (rows a and b at
// byte offsets va, vb, column offsets C i): the image value there is dC/dm,
// and dC/dtheta = (dC/dm + reg) beta m (1 - m) is written in its place, with
// the arithmetic of the separate chain kernel.
template <bool REG>
__device__ __forceinline__ void chain_task(float beta, const c32* d, int C, int va, int vb,
                                           __amdgpu_buffer_rsrc_t mrc,
                                           __amdgpu_buffer_rsrc_t rrc,
                                           __amdgpu_buffer_rsrc_t grc) {
#pragma unroll
  for (int i0 = 0; i0 < 32; i0 += 8) {
    __builtin_amdgcn_sched_barrier(0);
    float ma[8], mb[8], ra[8], rb[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int so = C * (i0 + k) * (int)sizeof(float);
      ma[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(mrc, va, so, 0));
      mb[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(mrc, vb, so, 0));
      if constexpr (REG) {
        ra[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rrc, va, so, 0));
        rb[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rrc, vb, so, 0));
      } else {
        ra[k] = rb[k] = 0.0f;
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int so = C * (i0 + k) * (int)sizeof(float);
      const float ga = (d[i0 + k].re + ra[k]) * beta * ma[k] * (1.0f - ma[k]);
      const float gb = (d[i0 + k].im + rb[k]) * beta * mb[k] * (1.0f - mb[k]);
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, ga), grc, va, so, 0);
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, gb), grc, vb, so, 0);
    }
  }
}

// MODE 0: the image.  MODE 1 (fidelity): instead of the image, the ILT
// fidelity of every pixel against the target (epi): per-work-group FP64
// partial sums of r^2 at epi.partial[img][p] and, when epi.G is set,
// G = dC/dI written where the image would be; image img is (target
// img / ncond, condition img % ncond).  MODE 2 (chain): the image is dC/dm of
// ILT row img, and dC/dtheta goes to row (ch.orows ? ch.orows[img] : img) of
// ch.grad (chain_task).
template <int TMIN, int TMAX, int MODE>
__global__ __launch_bounds__(256) void k_band_image(const float2* __restrict__ Ib, int n, int G,
                                                    int nimg, const float2* __restrict__ tw_n,
                                                    float* __restrict__ out, FidelityEpi epi,
                                                    ChainEpi ch) {
  constexpr bool FID = MODE == 1, CHAIN = MODE == 2;
  constexpr int F = 32, NT = TMAX - TMIN + 1;
  extern __shared__ c32 smem[];
  constexpr int UP = 32 * (-TMIN) + 1;  // U row: g = 0 .. 32 |TMIN|, zero beyond G
  const int P = n / F, C = n / F;
  c32* U = smem;            // F x UP
  c32* twh = U + F * UP;    // [h][c] = e^{2 pi i h c / n}, h < 32, c < C
  c32* twt = twh + F * C;   // [t - TMIN][c] = e^{2 pi i t c / C}
  const int L = blockIdx.x;
  const int slot = L / kXcds;
  const int p = slot % P;
  const int img = (slot / P) * kXcds + L % kXcds;
  if (img >= nimg) return;  // whole work-group, before any barrier
  const int tid = threadIdx.x;
  // Class-indexed tables: lane c of a phase-2 wave reads consecutive entries.
  for (int i = tid; i < F * C; i += 256) twh[i] = tw_lookup<+1>(tw_n, (i / C) * (i % C));
  for (int i = tid; i < NT * C; i += 256) {
    const int t = i / C + TMIN, c = i % C;
    twt[i] = tw_lookup<+1>(tw_n, (((F * t * c) % n) + n) % n);
  }
  for (int i = tid; i < F * (UP - (G + 1)); i += 256) {
    const int row = i / (UP - (G + 1)), col = G + 1 + i % (UP - (G + 1));
    U[row * UP + col] = c32{0.f, 0.f};
  }
  __syncthreads();
  const int nk = 2 * G + 1;
  // ---- phase 1: lanes l and l + 32 of wave v take column gx = 32 v + l; lane
  // half hi folds the bins h = 16 hi + h' (h' < 16) over every t.  One radix-2
  // step across the pair (a permlane32 swap) and an FFT_16 per lane then give
  // the rows j = 2 j' + hi of U. ----
  {
    constexpr int H = F / 2;
    const int lane = tid & 63, hi = lane >> 5;
    const int gx = 32 * (tid >> 6) + (lane & 31);
    const bool col = gx <= G;
    // lanes past the band fold a valid column and store nothing, so every
    // lane of a wave takes part in the swap
    const int gxa = col ? gx : G;
    const __amdgpu_buffer_rsrc_t br = box_rsrc(Ib + (size_t)img * nk * (G + 1), nk * (G + 1));
    int vo = (gxa + H * hi * (G + 1)) * (int)sizeof(float2);
    c32 u[H];
#pragma unroll
    for (int h = 0; h < H; ++h) u[h] = c32{0.f, 0.f};
    // The 16 band rows of fold t, this lane's column.
    auto load_block = [&](int t, c32* e) {
#pragma unroll
      for (int h = 0; h < H; ++h) {
        // band row H hi + h + F t at byte offset vo + ro
        const int ro = (h + F * t + G) * (G + 1) * (int)sizeof(float2);
        // Only the outermost folds can leave the band, and there the whole
        // offset goes through the lane's register with none in the scalar
        // offset: a row below the band gives a negative offset, which wraps
        // past the buffer's end, and a row above it an offset past the end;
        // the buffer range check returns zero for both.
        if (t == TMIN || t == TMAX) e[h] = load_c32(br, vo + ro, 0);
        else e[h] = load_c32(br, vo, ro);
      }
    };
    // One block of 16 loads in flight beside the accumulators (requesting
    // fold t + 1's rows before fold t is accumulated measured no faster).
#pragma unroll
    for (int t = TMIN; t <= TMAX; ++t) {
      __builtin_amdgcn_sched_barrier(0);
      c32 e[H];
      load_block(t, e);
      const c32 w = twt[(t - TMIN) * C + p];  // e^{2 pi i t p / P}
#pragma unroll
      for (int h = 0; h < H; ++h) u[h] = t == 0 ? cadd(u[h], e[h]) : cmac(u[h], e[h], w);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int h = 0; h < H; ++h) u[h] = cmul(u[h], twh[(H * hi + h) * C + p]);
    // radix-2 step of the inverse FFT_32 on x = bin h (lane hi = 0) and
    // bin h + 16 (lane hi = 1): lane 0 keeps x[h] + x[h + 16] (even rows),
    // lane 1 (x[h] - x[h + 16]) e^{2 pi i h / 32} (odd rows).  Swapping a = x
    // with b = x negated in the upper lanes leaves a = x[h] in every lane and
    // b = x[h + 16] (lower) or -x[h + 16] (upper): a + b is the butterfly.
    const float sg = hi ? -1.f : 1.f;
#pragma unroll
    for (int h = 0; h < H; ++h) {
      float ar = u[h].re, br = sg * u[h].re, ai = u[h].im, bi = sg * u[h].im;
      swap32(ar, br);
      swap32(ai, bi);
      u[h] = c32{ar + br, ai + bi};
    }
    tw_half_if<+1, H>(u, hi != 0);
    fft_reg<+1, H>(u);
    if (col) {
#pragma unroll
      for (int j = 0; j < H; ++j) U[(2 * j + hi) * UP + gx] = u[j];
    }
  }
  __syncthreads();
  // ---- pair packing, in place: Z[pr][g + GP] = z_pr[g] for |g| <= GP ----
  constexpr int GP = 32 * (-TMIN);
  constexpr int ZW = 2 * GP + 1;
  constexpr int NZ = (F / 2) * ZW;  // <= F * UP
  constexpr int PER = (NZ + 255) / 256;
  {
    c32 zv[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int idx = tid + 256 * k;
      zv[k] = c32{0.f, 0.f};
      if (idx < NZ) {
        const int pr = idx / ZW, g = idx % ZW - GP, ag = g < 0 ? -g : g;
        const c32 a = U[(2 * pr) * UP + ag], b = U[(2 * pr + 1) * UP + ag];
        zv[k] = g >= 0 ? c32{a.re - b.im, a.im + b.re} : c32{a.re + b.im, -a.im + b.re};
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int idx = tid + 256 * k;
      if (idx < NZ) U[idx] = zv[k];
    }
    __syncthreads();
  }
  // ---- phase 2 ----
  float* dst = FID ? epi.G : (CHAIN ? ch.grad : out);
  const int orow = CHAIN && ch.orows ? ch.orows[img] : img;
  const __amdgpu_buffer_rsrc_t orc = __builtin_amdgcn_make_buffer_rsrc(
      dst ? dst + (size_t)orow * n * n : nullptr, (short)0, n * n * (int)sizeof(float), 0x00020000);
  const __amdgpu_buffer_rsrc_t mrc = __builtin_amdgcn_make_buffer_rsrc(
      CHAIN ? const_cast<float*>(ch.m) + (size_t)img * n * n : nullptr, (short)0,
      n * n * (int)sizeof(float), 0x00020000);
  const __amdgpu_buffer_rsrc_t rrc = __builtin_amdgcn_make_buffer_rsrc(
      CHAIN && ch.reg ? const_cast<float*>(ch.reg) + (size_t)img * n * n : nullptr, (short)0,
      n * n * (int)sizeof(float), 0x00020000);
  const int tgt = FID ? img / epi.ncond : 0;
  const float dose = FID ? epi.dose[img % epi.ncond] : 1.0f;
  const __amdgpu_buffer_rsrc_t zrc = __builtin_amdgcn_make_buffer_rsrc(
      FID ? const_cast<float*>(epi.target) + (size_t)(epi.trows ? epi.trows[tgt] : tgt) * n * n
          : nullptr,
      (short)0, n * n * (int)sizeof(float), 0x00020000);
  double fid_acc = 0.0;
  // Lanes l and l + 32 of a wave share a task (row pair pr, classes c and
  // c + C/2, c < C/2): e^{2 pi i t (c + C/2) / C} = (-1)^t e^{2 pi i t c / C},
  // so the even-t and odd-t partial folds E, O give d_c = E + O and
  // d_{c+C/2} = E - O at once.  Lane l folds the bins h < 16, lane l + 32 the
  // bins h >= 16; one permlane32 swap per register then hands lane l all 32
  // bins of class c and lane l + 32 those of class c + C/2.  Every wave is
  // wholly active or wholly idle in each pass: the 8 C tasks (C = n/32 a
  // multiple of 4) fill whole waves of 32 tasks.
  constexpr int H = F / 2;
  const int lane = tid & 63, hi = lane >> 5;
  const int ntask = (F / 2) * (C / 2);
  for (int base = (tid >> 6) * 32; base < ntask; base += 128) {
    const int task = base + (lane & 31);
    const int pr = task / (C / 2);
    int c = task % (C / 2);
    const c32* zr = U + pr * ZW + GP + H * hi;  // zr[g - H hi], |g| <= GP
    c32 ev[H], od[H];
#pragma unroll
    for (int h = 0; h < H; ++h) ev[h] = od[h] = c32{0.f, 0.f};
#pragma unroll
    for (int t = TMIN; t <= TMAX; ++t) {
      __builtin_amdgcn_sched_barrier(0);
      const c32 w = twt[(t - TMIN) * C + c];  // e^{2 pi i t c / C}
#pragma unroll
      for (int h = 0; h < H; ++h) {
        // Z is zero beyond G, so every fold term is read unconditionally
        const c32 z = zr[h + F * t];
        if ((t & 1) == 0) ev[h] = t == 0 ? cadd(ev[h], z) : cmac(ev[h], z, w);
        else od[h] = cmac(od[h], z, w);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    c32 d[F];
#pragma unroll
    for (int h = 0; h < H; ++h) {
      // a = d_c (bins of this lane), b = d_{c + C/2}; after the swap the lower
      // lane holds d_c at bins h, h + 16 and the upper lane d_{c + C/2}
      c32 a = cadd(ev[h], od[h]), b = csub(ev[h], od[h]);
      swap32(a.re, b.re);
      swap32(a.im, b.im);
      d[h] = a;
      d[h + H] = b;
    }
    c += hi * (C / 2);  // this lane's class from here on
#pragma unroll
    for (int h = 0; h < F; ++h) d[h] = cmul(d[h], twh[h * C + c]);
    fft_reg<+1, F>(d);
    // Rows y = P (2 pr) + p and y + P, columns x = c + C i: vector buffer
    // stores with the column offset C i in an SGPR, marked non-temporal (cache
    // policy bit 1): the images stream out without evicting the next step's
    // inputs from L2 and the MALL.
    const int va = ((P * (2 * pr) + p) * n + c) * (int)sizeof(float);
    const int vb = va + P * n * (int)sizeof(float);
    if constexpr (CHAIN) {
      if (ch.reg) chain_task<true>(ch.beta, d, C, va, vb, mrc, rrc, orc);
      else chain_task<false>(ch.beta, d, C, va, vb, mrc, rrc, orc);
    } else if constexpr (!FID) {
#pragma unroll
      for (int i = 0; i < F; ++i) {
        const int so = C * i * (int)sizeof(float);
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, d[i].re), orc, va, so, 2);
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, d[i].im), orc, vb, so, 2);
      }
    } else {
      // the fidelity term of each pixel instead of the pixel (G = dC/dI in
      // place of the image when asked for); one branch per task on the cost
      // kind and on G, none per pixel
      if (epi.cost_kind == 0) {
        if (epi.G) fid_acc += fidelity_task<0, true>(epi, dose, d, C, va, vb, zrc, orc);
        else fid_acc += fidelity_task<0, false>(epi, dose, d, C, va, vb, zrc, orc);
      } else {
        if (epi.G) fid_acc += fidelity_task<1, true>(epi, dose, d, C, va, vb, zrc, orc);
        else fid_acc += fidelity_task<1, false>(epi, dose, d, C, va, vb, zrc, orc);
      }
    }
  }
  if constexpr (FID) {
    // work-group sum in a fixed order: lanes by a butterfly, then the waves
    __shared__ double wsum[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) fid_acc += __shfl_xor(fid_acc, off);
    if (lane == 0) wsum[tid >> 6] = fid_acc;
    __syncthreads();
    if (tid == 0) epi.partial[(size_t)img * P + p] = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
  }
}

}  // namespace huvk

// ===========================================================================
// Launchers
// ===========================================================================
namespace huvk {

int choose_coarse(int n, int B) {
  const int sizes[] = {32, 64, 128, 192, 256};
  for (int m : sizes)
    if (m >= 4 * B + 1 && m <= n) return m;
  return 0;
}

int coarse_classes(int m) { return m % 32 == 0 ? m / 32 : 0; }

int coarse_band(int m) { return (m - 1) / 4; }

#define HUV_LAUNCH(kern, grid, block, shmem, s, ...)                   \
  do {                                                                 \
    hipLaunchKernelGGL(kern, grid, block, shmem, s, __VA_ARGS__);      \
    return hipGetLastError();                                          \
  } while (0)

template <int N1, int N2>
static hipError_t rows_band_n(int B, int nclips, const float* masks, const float2* tw,
                              float2* tmp, hipStream_t s) {
  constexpr int TEAMS = 128 / N2 > 0 ? 128 / N2 : 1;
  const int npairs = N1 * N2 / 2;
  dim3 grid((npairs + TEAMS - 1) / TEAMS, nclips);
  HUV_LAUNCH((k_rows_band<N1, N2, TEAMS>), grid, dim3(TEAMS * N2), 0, s, masks, N1 * N2, B, tw,
             tmp);
}

int band_fft_default(int n) {
  switch (n) {
    case 256: return 16;
    case 512: return 16;
    case 1024: return 32;
    case 2048: return 32;
    default: return 0;
  }
}

int band_fft_options(int n, int* n1s) {
  switch (n) {
    case 256: n1s[0] = 8; n1s[1] = 16; return 2;
    case 512: n1s[0] = 8; n1s[1] = 16; return 2;
    case 1024: n1s[0] = 16; n1s[1] = 32; return 2;
    case 2048: n1s[0] = 32; return 1;
    default: return 0;
  }
}

// Dispatch over the supported four-step factorisations n = N1 x N2.
#define HUV_BAND_FFT_SWITCH(FN, ...)                                  \
  switch (n * 100 + (n1 ? n1 : band_fft_default(n))) {                \
    case 256 * 100 + 8: return FN<8, 32>(__VA_ARGS__);                 \
    case 256 * 100 + 16: return FN<16, 16>(__VA_ARGS__);               \
    case 512 * 100 + 8: return FN<8, 64>(__VA_ARGS__);                 \
    case 512 * 100 + 16: return FN<16, 32>(__VA_ARGS__);               \
    case 1024 * 100 + 16: return FN<16, 64>(__VA_ARGS__);              \
    case 1024 * 100 + 32: return FN<32, 32>(__VA_ARGS__);              \
    case 2048 * 100 + 32: return FN<32, 64>(__VA_ARGS__);              \
    default: return hipErrorInvalidValue;                              \
  }

hipError_t launch_rows_band(int n, int n1, int B, int nclips, const float* masks,
                            const float2* tw_n, float2* tmp, hipStream_t s) {
  HUV_BAND_FFT_SWITCH(rows_band_n, B, nclips, masks, tw_n, tmp, s)
}

template <int N1, int N2>
static hipError_t cols_band_n(int B, int BB, int nclips, const float2* tmp, const float2* tw_n,
                              const float* sinc_t, float2* Mq, float2* Mp, hipStream_t s) {
  constexpr int TEAMS = 128 / N2 > 0 ? 128 / N2 : 1;
  dim3 grid((BB + 1 + TEAMS - 1) / TEAMS, nclips);
  HUV_LAUNCH((k_cols_band<N1, N2, TEAMS>), grid, dim3(TEAMS * N2), 0, s, tmp, B, BB, tw_n, sinc_t,
             Mq, Mp);
}

hipError_t launch_cols_band(int n, int n1, int B, int BB, int nclips, const float2* tmp,
                            const float2* tw_n, const float* sinc_t, float2* Mq, float2* Mp,
                            hipStream_t s) {
  if (BB < B || 2 * BB + 1 > n) return hipErrorInvalidValue;
  HUV_BAND_FFT_SWITCH(cols_band_n, B, BB, nclips, tmp, tw_n, sinc_t, Mq, Mp, s)
}

template <int N1, int N2>
static hipError_t rows_band_c_n(int B, int nclips, const float2* masks, const float2* tw,
                                float2* tmp, hipStream_t s) {
  constexpr int TEAMS = 128 / N2 > 0 ? 128 / N2 : 1;
  dim3 grid((N1 * N2 + TEAMS - 1) / TEAMS, nclips);
  HUV_LAUNCH((k_rows_band_c<N1, N2, TEAMS>), grid, dim3(TEAMS * N2), 0, s, masks, B, tw, tmp);
}

hipError_t launch_rows_band_c(int n, int n1, int B, int nclips, const float2* masks,
                              const float2* tw_n, float2* tmp, hipStream_t s) {
  HUV_BAND_FFT_SWITCH(rows_band_c_n, B, nclips, masks, tw_n, tmp, s)
}

template <int N1, int N2>
static hipError_t cols_band_c_n(int B, int BB, int nclips, const float2* tmp, const float2* tw_n,
                                const float* sinc_t, float2* Mq, float2* Mp, hipStream_t s) {
  constexpr int TEAMS = 128 / N2 > 0 ? 128 / N2 : 1;
  dim3 grid((2 * BB + 1 + TEAMS - 1) / TEAMS, nclips);
  HUV_LAUNCH((k_cols_band_c<N1, N2, TEAMS>), grid, dim3(TEAMS * N2), 0, s, tmp, B, BB, tw_n,
             sinc_t, Mq, Mp);
}

hipError_t launch_cols_band_c(int n, int n1, int B, int BB, int nclips, const float2* tmp,
                              const float2* tw_n, const float* sinc_t, float2* Mq, float2* Mp,
                              hipStream_t s) {
  if (BB < B || 2 * BB + 1 > n) return hipErrorInvalidValue;
  HUV_BAND_FFT_SWITCH(cols_band_c_n, B, BB, nclips, tmp, tw_n, sinc_t, Mq, Mp, s)
}

template <int M, int XNB, int XS = 1>
static hipError_t fields_m(int nclips, int nplanes, int Kmax, int S, const float2* kern,
                           const float* lam, const float2* Mq, const float2* tw_m, float* coarse,
                           hipStream_t s) {
  using C = CoarseCfg<M, XS>;
  // the two-lane x pass needs every column in [-1, W)
  if constexpr (XS == 2 && !C::SPLIT_OK) return hipErrorInvalidValue;
  const size_t shmem = (size_t)C::LDS_C32 * sizeof(c32);
  auto kernel = k_socs_fields<M, XNB, XS>;
  // Once per instantiation (thread-safe static), never inside a graph capture.
  static const hipError_t attr = hipFuncSetAttribute(
      (const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
  if (attr != hipSuccess) return attr;
  const int groups = (nclips + kXcds - 1) / kXcds;
  HUV_LAUNCH(kernel, dim3(groups * kXcds * C::Q * S), dim3(C::THREADS), shmem, s, Mq, kern, lam,
             nclips, nplanes, Kmax, S, tw_m, coarse);
}

template <int M>
static hipError_t fields_x(int xnb, int xlanes, int nclips, int nplanes, int Kmax, int S,
                           const float2* kern, const float* lam, const float2* Mq,
                           const float2* tw_m, float* coarse, hipStream_t s) {
  if (!xlanes) xlanes = fields_lanes_default(M);
  if (!xnb) xnb = xlanes == 2 ? kFieldsXnbDefault2 : kFieldsXnbDefault;
  switch (xlanes * 100 + xnb) {
    case 104: return fields_m<M, 4>(nclips, nplanes, Kmax, S, kern, lam, Mq, tw_m, coarse, s);
    case 106: return fields_m<M, 6>(nclips, nplanes, Kmax, S, kern, lam, Mq, tw_m, coarse, s);
    case 108: return fields_m<M, 8>(nclips, nplanes, Kmax, S, kern, lam, Mq, tw_m, coarse, s);
    case 204: return fields_m<M, 4, 2>(nclips, nplanes, Kmax, S, kern, lam, Mq, tw_m, coarse, s);
    case 206: return fields_m<M, 6, 2>(nclips, nplanes, Kmax, S, kern, lam, Mq, tw_m, coarse, s);
    case 208: return fields_m<M, 8, 2>(nclips, nplanes, Kmax, S, kern, lam, Mq, tw_m, coarse, s);
    default: return hipErrorInvalidValue;
  }
}

int fields_lanes_default(int m) { return m >= 64 ? 2 : 1; }

bool fields_lanes_supported(int m, int xlanes) { return xlanes == 1 || (xlanes == 2 && m >= 64); }

int fields_splits(int m, int nclips, int Kmax) {
  const int Q = coarse_classes(m);
  if (Q <= 0 || Kmax <= 1) return 1;
  // small batches: deal the kernels over more work-groups (about 512 busy ones)
  const int wgs = nclips * Q;
  int S = (512 + wgs - 1) / wgs;
  if (S <= 1) return 1;
  if (S > Kmax) S = Kmax;
  const int ks = (Kmax + S - 1) / S;
  return (Kmax + ks - 1) / ks;  // every split non-empty
}

hipError_t launch_socs_fields(int m, int xnb, int xlanes, int nclips, int nplanes, int Kmax, int S,
                              const float2* kern, const float* lam, const float2* Mq,
                              const float2* tw_m, float* coarse, hipStream_t s) {
  if (S < 1 || S > Kmax) return hipErrorInvalidValue;
  switch (m) {
    case 32: return fields_x<32>(xnb, xlanes, nclips, nplanes, Kmax, S, kern, lam, Mq, tw_m, coarse, s);
    case 64: return fields_x<64>(xnb, xlanes, nclips, nplanes, Kmax, S, kern, lam, Mq, tw_m, coarse, s);
    case 128: return fields_x<128>(xnb, xlanes, nclips, nplanes, Kmax, S, kern, lam, Mq, tw_m, coarse, s);
    case 192: return fields_x<192>(xnb, xlanes, nclips, nplanes, Kmax, S, kern, lam, Mq, tw_m, coarse, s);
    case 256: return fields_x<256>(xnb, xlanes, nclips, nplanes, Kmax, S, kern, lam, Mq, tw_m, coarse, s);
    default: return hipErrorInvalidValue;
  }
}

template <int M>
static hipError_t coarse_rows_m(int B, int nimg, int S, const float* coarse, const float2* tw_m,
                                float2* tmp1, hipStream_t s) {
  constexpr int TEAMS = 256 / CoarseFft<M>::N2;
  dim3 grid((M / 2 + TEAMS - 1) / TEAMS, nimg);
  if (S > 1) {
    // fold the kernel-split partials into the first set (M^2 is a multiple of 4)
    const int64_t count4 = (int64_t)nimg * M * M / 4;
    const int blocks = (int)std::min<int64_t>((count4 + 255) / 256, 4096);
    hipLaunchKernelGGL(k_sum_splits, dim3(blocks), dim3(256), 0, s, const_cast<float*>(coarse), S,
                       count4);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  HUV_LAUNCH((k_coarse_rows<M, TEAMS>), grid, dim3(256), 0, s, coarse, B, tw_m, tmp1);
}

hipError_t launch_coarse_rows(int m, int B, int nimg, int S, const float* coarse,
                              const float2* tw_m, float2* tmp1, hipStream_t s) {
  if (S < 1) return hipErrorInvalidValue;
  switch (m) {
    case 32: return coarse_rows_m<32>(B, nimg, S, coarse, tw_m, tmp1, s);
    case 64: return coarse_rows_m<64>(B, nimg, S, coarse, tw_m, tmp1, s);
    case 128: return coarse_rows_m<128>(B, nimg, S, coarse, tw_m, tmp1, s);
    case 192: return coarse_rows_m<192>(B, nimg, S, coarse, tw_m, tmp1, s);
    case 256: return coarse_rows_m<256>(B, nimg, S, coarse, tw_m, tmp1, s);
    default: return hipErrorInvalidValue;
  }
}

template <int M, int THREADS>
static hipError_t coarse_fused_m(int B, int nimg, int S, const float* coarse, float flare,
                                 const float2* tw_m, float2* Ib, hipStream_t s) {
  if (S > 1) {
    const int64_t count4 = (int64_t)nimg * M * M / 4;
    const int blocks = (int)std::min<int64_t>((count4 + 255) / 256, 4096);
    hipLaunchKernelGGL(k_sum_splits, dim3(blocks), dim3(256), 0, s, const_cast<float*>(coarse), S,
                       count4);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  constexpr size_t shmem = (size_t)(M / 2) * (M + 1) * sizeof(c32);
  auto kern = k_coarse_fused<M, THREADS>;
  static const hipError_t attr = hipFuncSetAttribute(
      (const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
  if (attr != hipSuccess) return attr;
  HUV_LAUNCH(kern, dim3(nimg), dim3(THREADS), shmem, s, coarse, B, flare, tw_m, Ib);
}

hipError_t launch_coarse_fused(int m, int B, int nimg, int S, const float* coarse, float flare,
                               const float2* tw_m, float2* Ib, hipStream_t s) {
  if (S < 1 || 4 * B + 1 > m) return hipErrorInvalidValue;
  switch (m) {
    case 64: return coarse_fused_m<64, 512>(B, nimg, S, coarse, flare, tw_m, Ib, s);
    case 128: return coarse_fused_m<128, 512>(B, nimg, S, coarse, flare, tw_m, Ib, s);
    case 192: return coarse_fused_m<192, 512>(B, nimg, S, coarse, flare, tw_m, Ib, s);
    default: return hipErrorNotSupported;
  }
}

bool coarse_fused_supported(int m) { return m == 64 || m == 128 || m == 192; }

template <int M>
static hipError_t coarse_cols_m(int B, int nimg, float flare, const float2* tmp1,
                                const float2* tw_m, float2* Ib, hipStream_t s) {
  constexpr int TEAMS = 256 / CoarseFft<M>::N2;
  dim3 grid((2 * B + 1 + TEAMS - 1) / TEAMS, nimg);
  HUV_LAUNCH((k_coarse_cols<M, TEAMS>), grid, dim3(256), 0, s, tmp1, B, flare, tw_m, Ib);
}

hipError_t launch_coarse_cols(int m, int B, int nimg, float flare, const float2* tmp1,
                              const float2* tw_m, float2* Ib, hipStream_t s) {
  switch (m) {
    case 32: return coarse_cols_m<32>(B, nimg, flare, tmp1, tw_m, Ib, s);
    case 64: return coarse_cols_m<64>(B, nimg, flare, tmp1, tw_m, Ib, s);
    case 128: return coarse_cols_m<128>(B, nimg, flare, tmp1, tw_m, Ib, s);
    case 192: return coarse_cols_m<192>(B, nimg, flare, tmp1, tw_m, Ib, s);
    case 256: return coarse_cols_m<256>(B, nimg, flare, tmp1, tw_m, Ib, s);
    default: return hipErrorInvalidValue;
  }
}

template <int TMIN, int TMAX, int MODE>
static hipError_t band_image_t(int n, int G, int nimg, const float2* Ib, const float2* tw_n,
                               float* out, const FidelityEpi& epi, const ChainEpi& ch,
                               hipStream_t s) {
  constexpr int UP = 32 * (-TMIN) + 1, NT = TMAX - TMIN + 1;
  auto bytes = [](int nn) { return ((size_t)32 * UP + (size_t)(32 + NT) * (nn / 32)) * sizeof(c32); };
  const size_t shmem = bytes(n);
  auto kern = k_band_image<TMIN, TMAX, MODE>;
  // Sized for the largest grid (n = 2048), once.
  static const hipError_t attr = hipFuncSetAttribute(
      (const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes(2048));
  if (attr != hipSuccess) return attr;
  const int groups = (nimg + kXcds - 1) / kXcds;
  HUV_LAUNCH(kern, dim3(groups * kXcds * (n / 32)), dim3(256), shmem, s, Ib, n, G, nimg, tw_n,
             out, epi, ch);
}

template <int MODE>
static hipError_t band_image_g(int n, int G, int nimg, const float2* Ib, const float2* tw_n,
                               float* out, const FidelityEpi& epi, const ChainEpi& ch,
                               hipStream_t s) {
  if (n < 256 || n > 2048 || (n & (n - 1)) != 0 || G < 0 || G > 255 || 2 * G + 1 > n)
    return hipErrorInvalidValue;
  if (G <= 31) return band_image_t<-1, 0, MODE>(n, G, nimg, Ib, tw_n, out, epi, ch, s);
  if (G <= 63) return band_image_t<-2, 1, MODE>(n, G, nimg, Ib, tw_n, out, epi, ch, s);
  if (G <= 95) return band_image_t<-3, 2, MODE>(n, G, nimg, Ib, tw_n, out, epi, ch, s);
  if (G <= 127) return band_image_t<-4, 3, MODE>(n, G, nimg, Ib, tw_n, out, epi, ch, s);
  return hipErrorInvalidValue;
}

hipError_t launch_band_image(int n, int G, int nimg, const float2* Ib, const float2* tw_n,
                             float* out, hipStream_t s) {
  return band_image_g<0>(n, G, nimg, Ib, tw_n, out, FidelityEpi{}, ChainEpi{}, s);
}

hipError_t launch_band_image_fidelity(int n, int G, int nimg, const float2* Ib,
                                      const float2* tw_n, const FidelityEpi& epi, hipStream_t s) {
  if (!epi.target || !epi.dose || !epi.partial || epi.ncond < 1 || nimg % epi.ncond != 0)
    return hipErrorInvalidValue;
  return band_image_g<1>(n, G, nimg, Ib, tw_n, nullptr, epi, ChainEpi{}, s);
}

hipError_t launch_band_image_chain(int n, int G, int nimg, const float2* Ib, const float2* tw_n,
                                   const ChainEpi& ch, hipStream_t s) {
  if (!ch.m || !ch.grad) return hipErrorInvalidValue;
  return band_image_g<2>(n, G, nimg, Ib, tw_n, nullptr, FidelityEpi{}, ch, s);
}

}  // namespace huvk
