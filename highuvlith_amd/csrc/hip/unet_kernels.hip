// HIP kernels of the U-Net warm start (gfx950): BF16 activations and
// weights, FP32 accumulation on the matrix cores.
//
// 3x3 convolutions are implicit GEMMs on v_mfma_f32_32x32x16_bf16
// (D[32 x 32] += A[32 x 16] B[16 x 32]; lane l supplies A row l % 32 and
// B column l % 32 for k in 8 (l / 32) .. 8 (l / 32) + 7, and holds
// D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32] in accumulator r):
//   forward:  rows = output pixels, columns = output channels, k = (tap, ci);
//             NHWC activations and [co][tap][ci] weights give each lane's
//             8 consecutive k as one 16-byte load.
//   data gradient: the forward kernel on the ReLU-masked output gradient
//             with the rotated, transposed weights.
//   weight gradient: rows = output channels, columns = input channels of
//             one tap, k = pixels (16 consecutive x of one row), staged
//             channel-major in LDS from the NHWC tensors (rows a multiple of
//             64 wide) or read from NCHW copies (narrower rows); the +-1
//             column shift of a tap is a 2-byte funnel shift of aligned
//             words.  Split-K partials are summed in a fixed order.
#include <hip/hip_runtime.h>

#include "unet_kernels.hpp"

namespace huvk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ bf16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)0x7fc0;  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);                              // round to nearest even
  return (bf16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 ld8(const bf16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0)); }

// ---------------------------------------------------------------------------
// Forward 3x3 convolution.  Work-group = 4 waves along pixels; each wave
// computes BM x BN blocks of 32 pixels x 32 output channels.  The operands
// of step s + 1 are loaded before the products of step s are issued.
// ---------------------------------------------------------------------------
template <int BM, int BN>
__global__ __launch_bounds__(256) void k_conv3x3_fwd(const bf16_t* __restrict__ X, int N, int H,
                                                     int W, int Cin,
                                                     const bf16_t* __restrict__ Wt,
                                                     const float* __restrict__ bias, int Cout,
                                                     int relu, bf16_t* __restrict__ Y) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int half = lane >> 5, li = lane & 31;
  const int64_t NP = (int64_t)N * H * W;
  const int64_t pbase = ((int64_t)blockIdx.x * 4 + wave) * (32 * BM);
  const int cobase = blockIdx.y * 32 * BN;
  if (pbase >= NP) return;  // whole wave; no barriers in this kernel
  const int CB = Cin / 8, KB = 9 * CB, steps = (KB + 1) / 2;
  const int HW = H * W;
  int pn[BM], py[BM], px[BM];
  bool pv[BM];
#pragma unroll
  for (int bm = 0; bm < BM; ++bm) {
    const int64_t p = pbase + 32 * bm + li;
    pv[bm] = p < NP;
    const int64_t pp = pv[bm] ? p : 0;
    pn[bm] = (int)(pp / HW);
    const int rem = (int)(pp - (int64_t)pn[bm] * HW);
    py[bm] = rem / W;
    px[bm] = rem - py[bm] * W;
  }
  int co[BN];
#pragma unroll
  for (int bn = 0; bn < BN; ++bn) co[bn] = cobase + 32 * bn + li;

  auto load = [&](int s, bf16x8* a, bf16x8* b) {
    const int kb = 2 * s + half;
    const bool kv = kb < KB;
    const int tap = kv ? kb / CB : 0, cb = kv ? kb - (kb / CB) * CB : 0;
    const int dy = tap / 3 - 1, dx = tap - (tap / 3) * 3 - 1;
#pragma unroll
    for (int bm = 0; bm < BM; ++bm) {
      const int yy = py[bm] + dy, xx = px[bm] + dx;
      const bool v = kv && pv[bm] && yy >= 0 && yy < H && xx >= 0 && xx < W;
      a[bm] = v ? ld8(X + (((int64_t)pn[bm] * H + yy) * W + xx) * Cin + cb * 8) : zero8();
    }
#pragma unroll
    for (int bn = 0; bn < BN; ++bn) {
      const bool v = kv && co[bn] < Cout;
      b[bn] = v ? ld8(Wt + ((int64_t)co[bn] * 9 + tap) * Cin + cb * 8) : zero8();
    }
  };

  f32x16 acc[BM][BN];
#pragma unroll
  for (int bm = 0; bm < BM; ++bm)
#pragma unroll
    for (int bn = 0; bn < BN; ++bn)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[bm][bn][r] = 0.f;
  bf16x8 a[BM], b[BN], an[BM], bnx[BN];
  load(0, a, b);
  for (int s = 0; s < steps; ++s) {
    if (s + 1 < steps) load(s + 1, an, bnx);
#pragma unroll
    for (int bm = 0; bm < BM; ++bm)
#pragma unroll
      for (int bn = 0; bn < BN; ++bn)
        acc[bm][bn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[bm], b[bn], acc[bm][bn], 0, 0, 0);
#pragma unroll
    for (int bm = 0; bm < BM; ++bm) a[bm] = an[bm];
#pragma unroll
    for (int bn = 0; bn < BN; ++bn) b[bn] = bnx[bn];
  }
#pragma unroll
  for (int bn = 0; bn < BN; ++bn) {
    if (co[bn] >= Cout) continue;
    const float bv = bias ? bias[co[bn]] : 0.f;
#pragma unroll
    for (int bm = 0; bm < BM; ++bm)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int i = 8 * (r / 4) + 4 * half + (r % 4);
        const int64_t p = pbase + 32 * bm + i;
        if (p < NP) {
          float v = acc[bm][bn][r] + bv;
          if (relu) v = fmaxf(v, 0.f);
          Y[p * Cout + co[bn]] = f2bf(v);
        }
      }
  }
}

// ---------------------------------------------------------------------------
// Forward 3x3 convolution with the input and the weights staged in LDS.
// Work-group = 4 waves over an output tile of TH = 2 BM rows x 64 columns of
// one image; wave w takes columns x0 + 32 (w % 2) .. + 31 of rows
// y0 + w / 2 + 2 b (b < BM), and BN blocks of 32 output channels, so one
// weight fragment feeds BM products.  Input channels run in chunks of CC:
// the (TH + 2) x (64 + 2) x CC halo and the chunk's 9 x CC weights of the
// 32 BN output channels are loaded once per chunk (zero outside the image
// and beyond Cout) and read by all four waves from LDS; A = 32 pixels x 16 k
// per MFMA step (two 8-channel blocks of (tap, channel) index 2s and
// 2s + 1), B = 16 k x 32 output channels.  Needs W % 64 == 0 and
// H % TH == 0.
// ---------------------------------------------------------------------------
template <int BM, int BN, int CC>
__global__ __launch_bounds__(256) void k_conv3x3_fwd_lds(const bf16_t* __restrict__ X, int N,
                                                         int H, int W, int Cin,
                                                         const bf16_t* __restrict__ Wt,
                                                         const float* __restrict__ bias, int Cout,
                                                         int relu, bf16_t* __restrict__ Y) {
  constexpr int TH = 2 * BM, TW = 64, HR = TH + 2, HC = TW + 2;
  constexpr int PS = CC + 8;  // pixel stride in the tile (bf16), spreads the LDS banks
  constexpr int CB = CC / 8, KBC = 9 * CB, STEPS = (KBC + 1) / 2;
  constexpr int WS = 9 * CC + 8;  // output-channel stride of the staged weights (bf16)
  __shared__ __attribute__((aligned(16))) bf16_t tile[HR * HC * PS];
  __shared__ __attribute__((aligned(16))) bf16_t wts[32 * BN * WS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int half = lane >> 5, li = lane & 31;
  const int tilesX = W / TW, tilesY = H / TH;
  int blk = blockIdx.x;
  const int tx = blk % tilesX;
  blk /= tilesX;
  const int ty = blk % tilesY, n = blk / tilesY;
  const int x0 = tx * TW, y0 = ty * TH;
  const int cobase = blockIdx.y * 32 * BN;
  const int px = 32 * (wave & 1) + li;  // this lane's A row (pixel column) in the tile
  int co[BN];
#pragma unroll
  for (int bn = 0; bn < BN; ++bn) co[bn] = cobase + 32 * bn + li;
  f32x16 acc[BM][BN];
#pragma unroll
  for (int bm = 0; bm < BM; ++bm)
#pragma unroll
    for (int bn = 0; bn < BN; ++bn)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[bm][bn][r] = 0.f;

  for (int c0 = 0; c0 < Cin; c0 += CC) {
    // Every load of the chunk is issued before the first LDS store, so their
    // latencies overlap.
    constexpr int HALO = HR * HC * CB, HIT = (HALO + 255) / 256;
    constexpr int WGT = 32 * BN * KBC, WIT = (WGT + 255) / 256;
    uint4 hv[HIT], wv[WIT];
#pragma unroll
    for (int k = 0; k < HIT; ++k) {
      const int i = threadIdx.x + 256 * k;
      const int cb = i % CB, pix = i / CB;
      const int yy = y0 - 1 + pix / HC, xx = x0 - 1 + pix % HC;
      hv[k] = make_uint4(0, 0, 0, 0);
      if (i < HALO && yy >= 0 && yy < H && xx >= 0 && xx < W)
        hv[k] = *reinterpret_cast<const uint4*>(X + (((size_t)n * H + yy) * W + xx) * Cin + c0 +
                                                8 * cb);
    }
#pragma unroll
    for (int k = 0; k < WIT; ++k) {
      const int i = threadIdx.x + 256 * k;
      const int kb = i % KBC, col = i / KBC;  // kb = tap * CB + cb
      const int tap = kb / CB, cb = kb % CB, c = cobase + col;
      wv[k] = make_uint4(0, 0, 0, 0);
      if (i < WGT && c < Cout)
        wv[k] = *reinterpret_cast<const uint4*>(Wt + ((size_t)c * 9 + tap) * Cin + c0 + 8 * cb);
    }
    __syncthreads();  // the previous chunk has been consumed
#pragma unroll
    for (int k = 0; k < HIT; ++k) {
      const int i = threadIdx.x + 256 * k;
      if (i < HALO) *reinterpret_cast<uint4*>(tile + (i / CB) * PS + 8 * (i % CB)) = hv[k];
    }
#pragma unroll
    for (int k = 0; k < WIT; ++k) {
      const int i = threadIdx.x + 256 * k;
      if (i < WGT) *reinterpret_cast<uint4*>(wts + (i / KBC) * WS + 8 * (i % KBC)) = wv[k];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < STEPS; ++s) {
      const int kb = 2 * s + half;
      const bool kv = kb < KBC;
      const int tap = kv ? kb / CB : 0, cb = kv ? kb % CB : 0;
      const int dy = tap / 3, dx = tap % 3;
      bf16x8 bw[BN];
#pragma unroll
      for (int bn = 0; bn < BN; ++bn) {
        const bf16x8 w = ld8(wts + (32 * bn + li) * WS + 8 * (kv ? kb : 0));
        bw[bn] = kv ? w : zero8();
      }
#pragma unroll
      for (int bm = 0; bm < BM; ++bm) {
        const int py = (wave >> 1) + 2 * bm;
        const bf16x8 a0 = ld8(tile + ((py + dy) * HC + px + dx) * PS + 8 * cb);
        const bf16x8 a = kv ? a0 : zero8();
#pragma unroll
        for (int bn = 0; bn < BN; ++bn)
          acc[bm][bn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bw[bn], acc[bm][bn], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int bm = 0; bm < BM; ++bm) {
    const size_t prow = ((size_t)n * H + y0 + (wave >> 1) + 2 * bm) * W + x0 + 32 * (wave & 1);
#pragma unroll
    for (int bn = 0; bn < BN; ++bn) {
      if (co[bn] >= Cout) continue;
      const float bv = bias ? bias[co[bn]] : 0.f;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int i = 8 * (r / 4) + 4 * half + (r % 4);
        float v = acc[bm][bn][r] + bv;
        if (relu) v = fmaxf(v, 0.f);
        Y[(prow + i) * Cout + co[bn]] = f2bf(v);
      }
    }
  }
}

// Wr[ci][t][co] = W[co][8 - t][ci].
__global__ void k_conv_weight_rot(const bf16_t* __restrict__ Wt, int Cout, int Cin,
                                  bf16_t* __restrict__ Wr) {
  const int64_t total = (int64_t)Cout * 9 * Cin;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int ci = (int)(i % Cin);
    const int t = (int)((i / Cin) % 9);
    const int co = (int)(i / (9 * (int64_t)Cin));
    Wr[((int64_t)ci * 9 + (8 - t)) * Cout + co] = Wt[i];
  }
}

// Eight BF16 values as one 16-byte word.
union B8 {
  uint4 u;
  bf16_t h[8];
};

// ---------------------------------------------------------------------------
// NHWC -> NCHW through a 64-pixel x 32-channel LDS tile (HW % 64 == 0, C % 8
// == 0): each thread reads one 16-byte run of 8 channels of a pixel and
// writes one 16-byte run of 8 pixels of a channel.  MASK: the ReLU mask of
// the backward pass, dY zeroed (in place, and in the copy) where Y <= 0.
// ---------------------------------------------------------------------------
template <bool MASK>
__global__ __launch_bounds__(256) void k_transpose_t(bf16_t* __restrict__ dY,
                                                     const bf16_t* __restrict__ Y, int N, int HW,
                                                     int C, bf16_t* __restrict__ out) {
  constexpr int TP = 64, TC = 32, LP = TP + 8;  // padded pixel stride (16-byte aligned rows)
  __shared__ __align__(16) bf16_t tile[TC][LP];
  const int64_t p0 = (int64_t)blockIdx.x * TP;  // flat pixel over N * HW
  const int c0 = blockIdx.y * TC;
  {
    const int px = threadIdx.x >> 2, ch = 8 * (threadIdx.x & 3);
    if (c0 + ch < C) {
      const int64_t idx = (p0 + px) * C + c0 + ch;
      B8 v;
      v.u = *reinterpret_cast<const uint4*>(dY + idx);
      if (MASK && Y) {
        B8 y;
        y.u = *reinterpret_cast<const uint4*>(Y + idx);
        bool any = false;
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (bf2f(y.h[k]) <= 0.f) {
            v.h[k] = 0;
            any = true;
          }
        if (any) *reinterpret_cast<uint4*>(dY + idx) = v.u;
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) tile[ch + k][px] = v.h[k];
    }
  }
  __syncthreads();
  const int n = (int)(p0 / HW);
  const int hw0 = (int)(p0 - (int64_t)n * HW);
  const int c = threadIdx.x >> 3, pc = 8 * (threadIdx.x & 7);
  if (c0 + c < C)
    *reinterpret_cast<uint4*>(out + ((int64_t)n * C + c0 + c) * HW + hw0 + pc) =
        *reinterpret_cast<const uint4*>(&tile[c][pc]);
}

// ---------------------------------------------------------------------------
// Weight gradient (split K over rows).  Grid x = co-block groups (BM blocks
// of 32) * ci blocks * 3 kernel rows kh, grid y = splits.  Each wave loads a
// 16-pixel segment of dY (per co block) and of the input row y + kh - 1 once
// and feeds all three column taps kw from it (the -1 / +1 shifts are 2-byte
// funnel shifts), i.e. 3 BM products per BM + 1 operand loads.  The 4 waves
// take interleaved rows of the split and are reduced through LDS in a fixed
// order, one output tile at a time.
// ---------------------------------------------------------------------------
template <int BM>
__global__ __launch_bounds__(256) void k_conv3x3_wgrad(const bf16_t* __restrict__ dYT,
                                                       const bf16_t* __restrict__ XT, int N,
                                                       int H, int W, int Cin, int Cout,
                                                       int rows_per_split,
                                                       float* __restrict__ part) {
  __shared__ float red[4][32][33];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int half = lane >> 5, li = lane & 31;
  const int ciblocks = (Cin + 31) / 32;
  const int tile = blockIdx.x;
  const int kh = tile % 3, rest = tile / 3;
  const int cib = rest % ciblocks, cog = rest / ciblocks;
  const int dy = kh - 1;
  const int split = blockIdx.y;
  const int rows = N * H;
  const int r0 = split * rows_per_split;
  const int r1 = min(rows, r0 + rows_per_split);
  int coA[BM];
  bool av[BM];
#pragma unroll
  for (int bm = 0; bm < BM; ++bm) {
    coA[bm] = (cog * BM + bm) * 32 + li;
    av[bm] = coA[bm] < Cout;
  }
  const int ciB = cib * 32 + li;
  const bool bv = ciB < Cin;
  f32x16 acc[BM][3];
#pragma unroll
  for (int bm = 0; bm < BM; ++bm)
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[bm][k][r] = 0.f;
  for (int row = r0 + wave; row < r1; row += 4) {
    const int n = row / H, y = row - n * H;
    const int yy = y + dy;
    const bool rv = bv && yy >= 0 && yy < H;
    const bf16_t* arow[BM];
#pragma unroll
    for (int bm = 0; bm < BM; ++bm)
      arow[bm] = dYT + (((int64_t)n * Cout + (av[bm] ? coA[bm] : 0)) * H + y) * W;
    const bf16_t* brow = XT + (((int64_t)n * Cin + (bv ? ciB : 0)) * H + (rv ? yy : 0)) * W;
    // One 16-pixel segment: the dY rows (A), the input row (B) and its two
    // neighbouring words for the -1 / +1 column shifts.
    struct Seg {
      bf16x8 a[BM];
      uint4 w;
      uint32_t prev, next;
    };
    auto load = [&](int x0, Seg& g) {
      const int xs = x0 + 8 * half;
#pragma unroll
      for (int bm = 0; bm < BM; ++bm) g.a[bm] = av[bm] ? ld8(arow[bm] + xs) : zero8();
      g.w = make_uint4(0, 0, 0, 0);
      g.prev = 0u;
      g.next = 0u;
      if (rv) {
        g.w = *reinterpret_cast<const uint4*>(brow + xs);
        if (xs > 0) g.prev = *reinterpret_cast<const uint32_t*>(brow + xs - 2);
        if (xs + 8 < W) g.next = *reinterpret_cast<const uint32_t*>(brow + xs + 8);
      }
    };
    auto mma = [&](const Seg& g) {
      const uint4 w = g.w;
      // kw = 0: input column x - 1; kw = 1: x; kw = 2: x + 1
      const bf16x8 b0 = __builtin_bit_cast(
          bf16x8, make_uint4(__builtin_amdgcn_alignbit(w.x, g.prev, 16),
                             __builtin_amdgcn_alignbit(w.y, w.x, 16),
                             __builtin_amdgcn_alignbit(w.z, w.y, 16),
                             __builtin_amdgcn_alignbit(w.w, w.z, 16)));
      const bf16x8 b1 = __builtin_bit_cast(bf16x8, w);
      const bf16x8 b2 = __builtin_bit_cast(
          bf16x8, make_uint4(__builtin_amdgcn_alignbit(w.y, w.x, 16),
                             __builtin_amdgcn_alignbit(w.z, w.y, 16),
                             __builtin_amdgcn_alignbit(w.w, w.z, 16),
                             __builtin_amdgcn_alignbit(g.next, w.w, 16)));
#pragma unroll
      for (int bm = 0; bm < BM; ++bm) {
        acc[bm][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(g.a[bm], b0, acc[bm][0], 0, 0, 0);
        acc[bm][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(g.a[bm], b1, acc[bm][1], 0, 0, 0);
        acc[bm][2] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(g.a[bm], b2, acc[bm][2], 0, 0, 0);
      }
    };
    // Segments in groups of U, every load of a group issued before its
    // products, so the loads' latencies overlap.
    constexpr int U = 4;
    int x0 = 0;
    for (; x0 + 16 * U <= W; x0 += 16 * U) {
      Seg g[U];
#pragma unroll
      for (int u = 0; u < U; ++u) load(x0 + 16 * u, g[u]);
#pragma unroll
      for (int u = 0; u < U; ++u) mma(g[u]);
    }
    for (; x0 < W; x0 += 16) {
      Seg g;
      load(x0, g);
      mma(g);
    }
  }
#pragma unroll
  for (int bm = 0; bm < BM; ++bm)
#pragma unroll
    for (int kw = 0; kw < 3; ++kw) {
#pragma unroll
      for (int r = 0; r < 16; ++r) red[wave][8 * (r / 4) + 4 * half + (r % 4)][li] = acc[bm][kw][r];
      __syncthreads();
      for (int e = threadIdx.x; e < 32 * 32; e += 256) {
        const int i = e / 32, j = e % 32;
        const float v = ((red[0][i][j] + red[1][i][j]) + red[2][i][j]) + red[3][i][j];
        const int co = (cog * BM + bm) * 32 + i, ci = cib * 32 + j;
        if (co < Cout && ci < Cin)
          part[(((int64_t)split * Cout + co) * 9 + kh * 3 + kw) * Cin + ci] = v;
      }
      __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Weight and bias gradients read straight from the NHWC tensors.  Work-group
// = 3 waves on one block of 32 output channels x one block of 32 input
// channels x a contiguous range of image rows (split K).  The 64-pixel
// column segments are the outer loop and the rows the inner one: for each
// row y, dY[y][seg][co] and the input row y + 1 (columns seg - 1 .. seg + 64)
// are staged in LDS channel-major (one 16-byte word holds 8 consecutive
// pixels of a channel); the input rows y - 1 and y stay from the previous
// rows in a ring of three (all three are loaded at a segment's first row and
// where an image ends).  Wave ky takes the kernel row ky: for every 16
// pixels, A = dY (32 co x 16 px) and B(kx) = input row y + ky - 1 shifted by
// kx - 1 pixels (the one- and two-pixel shifts are funnel shifts of the
// aligned word and the next), acc[kx] += A B, so each wave owns its three
// taps outright.  Wave 0 of the work-groups of input block 0 also sums dY
// over the pixels (the bias gradient).  part[split][co][tap][ci],
// bpart[split][co].  Needs W % 64 == 0, Cin % 8 == 0, Cout % 8 == 0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(192) void k_conv3x3_wgrad_nhwc(
    const bf16_t* __restrict__ dY, const bf16_t* __restrict__ X, int N, int H, int W, int Cin,
    int Cout, int rows_per_split, float* __restrict__ part, float* __restrict__ bpart) {
  constexpr int SEG = 64, LP = SEG + 8;  // staged rows: 72 pixels (66 used), 16-byte aligned
  constexpr int NT = 192;
  constexpr int DT = SEG * 4, DI = (DT + NT - 1) / NT;            // dY: 64 px x 4 groups
  constexpr int RT = (SEG + 2) * 4;                               // one input row: 66 px x 4
  constexpr int XT = 3 * RT;                                      // three input rows
  __shared__ __align__(16) bf16_t dyT[32][LP];
  __shared__ __align__(16) bf16_t xT[3][32][LP];  // ring: input row yy in slot yy % 3
  const int tid = threadIdx.x;
  const int lane = tid & 63, ky = tid >> 6;
  const int half = lane >> 5, li = lane & 31;
  const int ciblocks = (Cin + 31) / 32;
  const int cib = blockIdx.x % ciblocks, cob = blockIdx.x / ciblocks;
  const int co0 = 32 * cob, ci0 = 32 * cib;
  const int split = blockIdx.y;
  const int rows = N * H;
  const int r0 = split * rows_per_split, r1 = min(rows, r0 + rows_per_split);
  const bool bias = bpart != nullptr && cib == 0 && ky == 0;
  f32x16 acc[3];
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[b][r] = 0.f;
  float bsum = 0.f;
  for (int x0 = 0; x0 < W; x0 += SEG) {
    for (int row = r0; row < r1; ++row) {
      const int n = row / H, y = row - n * H;
      // rows y - 1 .. y + 1 at a segment's first row or an image's first row
      // (a plain loop), else row y + 1 only (every load of the row issued
      // before the barrier)
      const bool all3 = row == r0 || y == 0;
      const int yfirst = all3 ? y - 1 : y + 1, nrow = all3 ? 3 : 1;
      auto x_load = [&](int t) -> uint4 {
        const int r = t / RT, rem = t % RT;
        const int jj = rem >> 2, cg = rem & 3;
        const int yy = yfirst + r, xx = x0 + jj - 1;
        if (r < nrow && yy >= 0 && yy < H && xx >= 0 && xx < W && ci0 + 8 * cg < Cin)
          return *reinterpret_cast<const uint4*>(X + (((size_t)n * H + yy) * W + xx) * Cin + ci0 +
                                                 8 * cg);
        return make_uint4(0, 0, 0, 0);
      };
      auto x_store = [&](int t, const B8& v) {
        const int r = t / RT, rem = t % RT;
        const int slot = (yfirst + r + 3) % 3;
#pragma unroll
        for (int e = 0; e < 8; ++e) xT[slot][8 * (rem & 3) + e][rem >> 2] = v.h[e];
      };
      B8 dv[DI];
#pragma unroll
      for (int k = 0; k < DI; ++k) {
        const int t = tid + NT * k, px = t >> 2, cg = t & 3;
        dv[k].u = make_uint4(0, 0, 0, 0);
        if (t < DT && co0 + 8 * cg < Cout)
          dv[k].u = *reinterpret_cast<const uint4*>(
              dY + (((size_t)n * H + y) * W + x0 + px) * Cout + co0 + 8 * cg);
      }
      constexpr int X1 = (RT + NT - 1) / NT;
      B8 xv[X1];
      if (!all3) {
#pragma unroll
        for (int k = 0; k < X1; ++k) {
          const int t = tid + NT * k;
          xv[k].u = t < RT ? x_load(t) : make_uint4(0, 0, 0, 0);
        }
      }
      __syncthreads();  // the previous row has been read
#pragma unroll
      for (int k = 0; k < DI; ++k) {
        const int t = tid + NT * k;
        if (t < DT) {
#pragma unroll
          for (int e = 0; e < 8; ++e) dyT[8 * (t & 3) + e][t >> 2] = dv[k].h[e];
        }
      }
      if (all3) {
        for (int t = tid; t < XT; t += NT) {
          B8 v;
          v.u = x_load(t);
          x_store(t, v);
        }
      } else {
#pragma unroll
        for (int k = 0; k < X1; ++k) {
          const int t = tid + NT * k;
          if (t < RT) x_store(t, xv[k]);
        }
      }
      __syncthreads();
      const int slot = (y + ky + 2) % 3;  // input row y + ky - 1
#pragma unroll
      for (int ks = 0; ks < SEG / 16; ++ks) {
        const int j0 = 16 * ks + 8 * half;  // this lane's 8 pixels
        const uint4 au = *reinterpret_cast<const uint4*>(&dyT[li][j0]);
        const bf16x8 a = __builtin_bit_cast(bf16x8, au);
        if (bias) {
          B8 ab;
          ab.u = au;
#pragma unroll
          for (int e = 0; e < 8; ++e) bsum += bf2f(ab.h[e]);
        }
        const uint4 w = *reinterpret_cast<const uint4*>(&xT[slot][li][j0]);
        const uint32_t nx = *reinterpret_cast<const uint32_t*>(&xT[slot][li][j0 + 8]);
        const bf16x8 b0 = __builtin_bit_cast(bf16x8, w);
        const bf16x8 b1 = __builtin_bit_cast(
            bf16x8, make_uint4(__builtin_amdgcn_alignbit(w.y, w.x, 16),
                               __builtin_amdgcn_alignbit(w.z, w.y, 16),
                               __builtin_amdgcn_alignbit(w.w, w.z, 16),
                               __builtin_amdgcn_alignbit(nx, w.w, 16)));
        const bf16x8 b2 = __builtin_bit_cast(bf16x8, make_uint4(w.y, w.z, w.w, nx));
        acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b0, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b1, acc[1], 0, 0, 0);
        acc[2] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b2, acc[2], 0, 0, 0);
      }
    }
  }
  // lane (li, half) holds rows 8 (r / 4) + 4 half + r % 4 (output channels) of
  // column li (input channel) of its wave's three taps
  const int ci = ci0 + li;
#pragma unroll
  for (int kx = 0; kx < 3; ++kx)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int co = co0 + 8 * (r / 4) + 4 * half + (r % 4);
      if (co < Cout && ci < Cin)
        part[(((size_t)split * Cout + co) * 9 + 3 * ky + kx) * Cin + ci] = acc[kx][r];
    }
  if (bias) {
    bsum += __shfl_xor(bsum, 32);  // the two pixel halves of channel li
    if (half == 0 && co0 + li < Cout) bpart[(size_t)split * Cout + co0 + li] = bsum;
  }
}

// dY *= (Y > 0), in place (NHWC, 8 channels per thread).
__global__ void k_relu_mask(bf16_t* __restrict__ dY, const bf16_t* __restrict__ Y,
                            int64_t count8) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count8;
       i += (int64_t)gridDim.x * blockDim.x) {
    B8 v, yv;
    v.u = reinterpret_cast<const uint4*>(dY)[i];
    yv.u = reinterpret_cast<const uint4*>(Y)[i];
    bool any = false;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (bf2f(yv.h[k]) <= 0.f) {
        v.h[k] = 0;
        any = true;
      }
    if (any) reinterpret_cast<uint4*>(dY)[i] = v.u;
  }
}

// out[i] = sum over k of part[k][i]: 32 outputs per work-group, the splits
// dealt round-robin over 32 groups of 32 lanes (each summed in increasing k),
// then the 32 group sums added in group order, so the result does not depend
// on scheduling.
__global__ __launch_bounds__(1024) void k_sum_splits(const float* __restrict__ part, int splits,
                                                     int64_t count, float* __restrict__ out) {
  __shared__ float sh[32][33];
  const int j = threadIdx.x & 31, g = threadIdx.x >> 5;
  const int64_t i = (int64_t)blockIdx.x * 32 + j;
  float s = 0.f;
  if (i < count)
    for (int k = g; k < splits; k += 32) s += part[(int64_t)k * count + i];
  sh[g][j] = s;
  __syncthreads();
  if (g == 0 && i < count) {
    float t = 0.f;
#pragma unroll
    for (int q = 0; q < 32; ++q) t += sh[q][j];
    out[i] = t;
  }
}

// part[s][c] = sum of dYT[n][c][hw] over the s-th of `splits` contiguous
// chunks of the flat (n, hw) range (fixed order within a chunk).
__global__ __launch_bounds__(256) void k_channel_partial(const bf16_t* __restrict__ dYT, int N,
                                                         int HW, int C, int64_t chunk,
                                                         float* __restrict__ part) {
  __shared__ float sh[256];
  const int c = blockIdx.x, sp = blockIdx.y;
  const int64_t total = (int64_t)N * HW;
  const int64_t i0 = (int64_t)sp * chunk, i1 = min(total, i0 + chunk);
  float s = 0.f;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
    const int64_t n = i / HW, hw = i - n * HW;
    s += bf2f(dYT[(n * C + c) * HW + hw]);
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)sp * C + c] = sh[0];
}

// ---------------------------------------------------------------------------
// Pooling, upsampling + concatenation, head, loss, optimiser.
// ---------------------------------------------------------------------------
// Element-wise kernels on NHWC tensors with channel counts that are multiples
// of 8: one thread per (pixel, 8-channel group), 16-byte loads and stores.
__device__ __forceinline__ B8 ld_b8(const bf16_t* p) {
  B8 r;
  r.u = *reinterpret_cast<const uint4*>(p);
  return r;
}
__device__ __forceinline__ void st_b8(bf16_t* p, const B8& v) {
  *reinterpret_cast<uint4*>(p) = v.u;
}

__global__ void k_maxpool_fwd(const bf16_t* __restrict__ X, int N, int H, int W, int C,
                              bf16_t* __restrict__ Y) {
  const int Ho = H / 2, Wo = W / 2, G = C / 8;
  const int total = N * Ho * Wo * G;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const int g = i % G, q = i / G;
    const int xo = q % Wo, t = q / Wo, yo = t % Ho, n = t / Ho;
    const bf16_t* b = X + (((size_t)n * H + 2 * yo) * W + 2 * xo) * C + 8 * g;
    const B8 v0 = ld_b8(b), v1 = ld_b8(b + C), v2 = ld_b8(b + (size_t)W * C),
             v3 = ld_b8(b + (size_t)W * C + C);
    B8 o;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      // the maximum is one of the inputs: keep its bits (no rounding)
      bf16_t m = v0.h[e];
      if (bf2f(v1.h[e]) > bf2f(m)) m = v1.h[e];
      if (bf2f(v2.h[e]) > bf2f(m)) m = v2.h[e];
      if (bf2f(v3.h[e]) > bf2f(m)) m = v3.h[e];
      o.h[e] = m;
    }
    st_b8(Y + (size_t)q * C + 8 * g, o);
  }
}

// dX: dY routed to the first maximum of each 2 x 2 window, zero elsewhere.
__global__ void k_maxpool_bwd(const bf16_t* __restrict__ X, const bf16_t* __restrict__ dY, int N,
                              int H, int W, int C, bf16_t* __restrict__ dX) {
  const int Ho = H / 2, Wo = W / 2, G = C / 8;
  const int total = N * Ho * Wo * G;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const int g = i % G, q = i / G;
    const int xo = q % Wo, t = q / Wo, yo = t % Ho, n = t / Ho;
    const size_t base = (((size_t)n * H + 2 * yo) * W + 2 * xo) * C + 8 * g;
    const size_t o[4] = {0, (size_t)C, (size_t)W * C, (size_t)W * C + C};
    B8 v[4], d[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = ld_b8(X + base + o[k]);
    const B8 dy = ld_b8(dY + (size_t)q * C + 8 * g);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      int arg = 0;
      float m = bf2f(v[0].h[e]);
#pragma unroll
      for (int k = 1; k < 4; ++k) {
        const float x = bf2f(v[k].h[e]);
        if (x > m) { m = x; arg = k; }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) d[k].h[e] = k == arg ? dy.h[e] : (bf16_t)0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) st_b8(dX + base + o[k], d[k]);
  }
}

// cat = [nearest 2x upsample of U (Cu channels), S (Cs channels)]
__global__ void k_upcat_fwd(const bf16_t* __restrict__ U, int Cu, const bf16_t* __restrict__ S,
                            int Cs, int N, int H, int W, bf16_t* __restrict__ cat) {
  const int Ct = Cu + Cs, G = Ct / 8, Gu = Cu / 8;
  const int total = N * H * W * G;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const int g = i % G, p = i / G;
    B8 v;
    if (g < Gu) {
      const int x = p % W, t = p / W, y = t % H, n = t / H;
      v = ld_b8(U + (((size_t)n * (H / 2) + y / 2) * (W / 2) + x / 2) * Cu + 8 * g);
    } else {
      v = ld_b8(S + (size_t)p * Cs + 8 * (g - Gu));
    }
    st_b8(cat + (size_t)p * Ct + 8 * g, v);
  }
}

// dU = sum of dcat's first Cu channels over each 2 x 2 block
__global__ void k_upcat_bwd_u(const bf16_t* __restrict__ dcat, int Cu, int Cs, int N, int H,
                              int W, bf16_t* __restrict__ dU) {
  const int Ct = Cu + Cs, Hh = H / 2, Wh = W / 2, G = Cu / 8;
  const int total = N * Hh * Wh * G;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const int g = i % G, q = i / G;
    const int x = q % Wh, t = q / Wh, y = t % Hh, n = t / Hh;
    const size_t b = (((size_t)n * H + 2 * y) * W + 2 * x) * Ct + 8 * g;
    const B8 a0 = ld_b8(dcat + b), a1 = ld_b8(dcat + b + Ct), a2 = ld_b8(dcat + b + (size_t)W * Ct),
             a3 = ld_b8(dcat + b + (size_t)W * Ct + Ct);
    B8 o;
#pragma unroll
    for (int e = 0; e < 8; ++e)
      o.h[e] = f2bf((bf2f(a0.h[e]) + bf2f(a1.h[e])) + (bf2f(a2.h[e]) + bf2f(a3.h[e])));
    st_b8(dU + (size_t)q * Cu + 8 * g, o);
  }
}

// dS (+)= dcat's last Cs channels
__global__ void k_upcat_bwd_s(const bf16_t* __restrict__ dcat, int Cu, int Cs, int64_t P,
                              bf16_t* __restrict__ dS, int accumulate) {
  const int Ct = Cu + Cs, G = Cs / 8;
  const int64_t total = P * G;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int g = (int)(i % G);
    const int64_t p = i / G;
    const B8 v = ld_b8(dcat + (size_t)p * Ct + Cu + 8 * g);
    bf16_t* d = dS + (size_t)p * Cs + 8 * g;
    if (accumulate) {
      const B8 w = ld_b8(d);
      B8 o;
#pragma unroll
      for (int e = 0; e < 8; ++e) o.h[e] = f2bf(bf2f(w.h[e]) + bf2f(v.h[e]));
      st_b8(d, o);
    } else {
      st_b8(d, v);
    }
  }
}

template <int C>
__global__ void k_head_fwd(const bf16_t* __restrict__ Y, int64_t P, const float* __restrict__ w,
                           const float* __restrict__ b, float* __restrict__ out) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < P;
       p += (int64_t)gridDim.x * blockDim.x) {
    float s = b[0];
#pragma unroll
    for (int c = 0; c < C; ++c) s += bf2f(Y[p * C + c]) * w[c];
    out[p] = s;
  }
}

// Block-partial dw (C values) and db (1 value) per block, fixed order.
constexpr int kHeadThreads = 256;
template <int C>
__global__ __launch_bounds__(kHeadThreads) void k_head_bwd(
    const bf16_t* __restrict__ Y, const float* __restrict__ dout, int64_t P,
    const float* __restrict__ w, bf16_t* __restrict__ dY, float* __restrict__ part) {
  __shared__ float sh[kHeadThreads];
  float acc[C + 1];
#pragma unroll
  for (int c = 0; c <= C; ++c) acc[c] = 0.f;
  for (int64_t p = blockIdx.x * (int64_t)kHeadThreads + threadIdx.x; p < P;
       p += (int64_t)gridDim.x * kHeadThreads) {
    const float g = dout[p];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      dY[p * C + c] = f2bf(g * w[c]);
      acc[c] += g * bf2f(Y[p * C + c]);
    }
    acc[C] += g;
  }
#pragma unroll
  for (int c = 0; c <= C; ++c) {
    sh[threadIdx.x] = acc[c];
    __syncthreads();
    for (int s = kHeadThreads / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * (C + 1) + c] = sh[0];
    __syncthreads();
  }
}

constexpr int kMseThreads = 256;
__global__ __launch_bounds__(kMseThreads) void k_mse(const float* __restrict__ out,
                                                     const float* __restrict__ target, int64_t P,
                                                     float* __restrict__ dout,
                                                     double* __restrict__ part) {
  __shared__ double sh[kMseThreads];
  double s = 0.0;
  const float sc = 2.0f / (float)P;
  for (int64_t p = blockIdx.x * (int64_t)kMseThreads + threadIdx.x; p < P;
       p += (int64_t)gridDim.x * kMseThreads) {
    const float d = out[p] - target[p];
    s += (double)d * d;
    dout[p] = sc * d;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kMseThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_pack_input(const float* __restrict__ in, int64_t P, bf16_t* __restrict__ X8) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < P;
       p += (int64_t)gridDim.x * blockDim.x) {
    uint4 v = make_uint4((uint32_t)f2bf(in[p]), 0u, 0u, 0u);
    *reinterpret_cast<uint4*>(X8 + p * 8) = v;
  }
}

__global__ void k_add_bf16(bf16_t* __restrict__ dS, const bf16_t* __restrict__ dX, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dS[i] = f2bf(bf2f(dS[i]) + bf2f(dX[i]));
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

__global__ void k_to_bf16(const float* __restrict__ x, int64_t n, bf16_t* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = f2bf(x[i]);
}

// ===========================================================================
// Launchers
// ===========================================================================
static dim3 grid1d(int64_t n, int threads = 256) {
  int64_t b = (n + threads - 1) / threads;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

hipError_t launch_conv3x3_fwd(const bf16_t* X, int N, int H, int W, int Cin, const bf16_t* Wt,
                              const float* bias, int Cout, int relu, bf16_t* Y, hipStream_t s) {
  if (Cin % 8 != 0 || Cout < 1 || N < 1 || H < 1 || W < 1) return hipErrorInvalidValue;
  const int64_t NP = (int64_t)N * H * W;
  if (W % 64 == 0 && H % 4 == 0 && Cin % 8 == 0 && (Cin <= 16 || Cin % 16 == 0) &&
      (int64_t)N * (H / 4) * (W / 64) < (1ll << 31)) {
    // tiles of 4 rows x 64 columns: BM = 2 pixel blocks per wave
    dim3 grid((unsigned)((int64_t)N * (H / 4) * (W / 64)),
              (unsigned)((Cout + (Cout <= 32 ? 31 : 63)) / (Cout <= 32 ? 32 : 64)));
#define HUV_CONV_LDS(BN_, CC_)                                                                    \
  hipLaunchKernelGGL((k_conv3x3_fwd_lds<2, BN_, CC_>), grid, dim3(256), 0, s, X, N, H, W, Cin, Wt, \
                     bias, Cout, relu, Y)
    if (Cout <= 32) {
      if (Cin == 8) HUV_CONV_LDS(1, 8);
      else if (Cin % 32 != 0) HUV_CONV_LDS(1, 16);
      else HUV_CONV_LDS(1, 32);
    } else {
      if (Cin == 8) HUV_CONV_LDS(2, 8);
      else if (Cin % 32 != 0) HUV_CONV_LDS(2, 16);
      else HUV_CONV_LDS(2, 32);
    }
#undef HUV_CONV_LDS
    return hipGetLastError();
  }
  constexpr int BM = 2;
  if (Cout <= 32) {
    dim3 grid((unsigned)((NP + 4 * 32 * BM - 1) / (4 * 32 * BM)), (unsigned)((Cout + 31) / 32));
    hipLaunchKernelGGL((k_conv3x3_fwd<BM, 1>), grid, dim3(256), 0, s, X, N, H, W, Cin, Wt, bias,
                       Cout, relu, Y);
  } else {
    dim3 grid((unsigned)((NP + 4 * 32 * BM - 1) / (4 * 32 * BM)), (unsigned)((Cout + 63) / 64));
    hipLaunchKernelGGL((k_conv3x3_fwd<BM, 2>), grid, dim3(256), 0, s, X, N, H, W, Cin, Wt, bias,
                       Cout, relu, Y);
  }
  return hipGetLastError();
}

hipError_t launch_conv_weight_rot(const bf16_t* Wt, int Cout, int Cin, bf16_t* Wr,
                                  hipStream_t s) {
  hipLaunchKernelGGL(k_conv_weight_rot, grid1d((int64_t)Cout * 9 * Cin), dim3(256), 0, s, Wt,
                     Cout, Cin, Wr);
  return hipGetLastError();
}

hipError_t launch_nhwc_to_nchw(const bf16_t* X, int N, int H, int W, int C, bf16_t* XT,
                               hipStream_t s) {
  if ((H * W) % 64 != 0 || C % 8 != 0) return hipErrorInvalidValue;
  dim3 grid((unsigned)((int64_t)N * H * W / 64), (unsigned)((C + 31) / 32));
  hipLaunchKernelGGL((k_transpose_t<false>), grid, dim3(256), 0, s, const_cast<bf16_t*>(X),
                     nullptr, N, H * W, C, XT);
  return hipGetLastError();
}

hipError_t launch_relu_mask_t(bf16_t* dY, const bf16_t* Y, int N, int H, int W, int C,
                              bf16_t* dYT, hipStream_t s) {
  if ((H * W) % 64 != 0 || C % 8 != 0) return hipErrorInvalidValue;
  dim3 grid((unsigned)((int64_t)N * H * W / 64), (unsigned)((C + 31) / 32));
  hipLaunchKernelGGL((k_transpose_t<true>), grid, dim3(256), 0, s, dY, Y, N, H * W, C, dYT);
  return hipGetLastError();
}

hipError_t launch_conv3x3_wgrad(const bf16_t* dYT, const bf16_t* XT, int N, int H, int W,
                                int Cin, int Cout, int splits, float* part, hipStream_t s) {
  if (W % 16 != 0 || Cin % 8 != 0 || splits < 1) return hipErrorInvalidValue;
  const int rows = N * H;
  const int rps = (rows + splits - 1) / splits;
  const int coblocks = (Cout + 31) / 32, ciblocks = (Cin + 31) / 32;
  if (coblocks >= 2) {
    const int tiles = ((coblocks + 1) / 2) * ciblocks * 3;
    hipLaunchKernelGGL(k_conv3x3_wgrad<2>, dim3(tiles, splits), dim3(256), 0, s, dYT, XT, N, H,
                       W, Cin, Cout, rps, part);
  } else {
    const int tiles = coblocks * ciblocks * 3;
    hipLaunchKernelGGL(k_conv3x3_wgrad<1>, dim3(tiles, splits), dim3(256), 0, s, dYT, XT, N, H,
                       W, Cin, Cout, rps, part);
  }
  return hipGetLastError();
}

hipError_t launch_conv3x3_wgrad_nhwc(const bf16_t* dY, const bf16_t* X, int N, int H, int W,
                                     int Cin, int Cout, int splits, float* part, float* bpart,
                                     hipStream_t s) {
  if (W % 64 != 0 || Cin % 8 != 0 || Cout % 8 != 0 || splits < 1 || splits > N * H)
    return hipErrorInvalidValue;
  const int rps = (N * H + splits - 1) / splits;
  const int blocks = ((Cout + 31) / 32) * ((Cin + 31) / 32);
  hipLaunchKernelGGL(k_conv3x3_wgrad_nhwc, dim3(blocks, splits), dim3(192), 0, s, dY, X, N, H, W,
                     Cin, Cout, rps, part, bpart);
  return hipGetLastError();
}

hipError_t launch_relu_mask(bf16_t* dY, const bf16_t* Y, int N, int H, int W, int C,
                            hipStream_t s) {
  if (C % 8 != 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_relu_mask, grid1d((int64_t)N * H * W * C / 8), dim3(256), 0, s, dY, Y,
                     (int64_t)N * H * W * C / 8);
  return hipGetLastError();
}

hipError_t launch_sum_splits(const float* part, int splits, int64_t count, float* out,
                             hipStream_t s) {
  if (count <= 0 || splits < 1 || (count + 31) / 32 >= (1ll << 31)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_sum_splits, dim3((unsigned)((count + 31) / 32)), dim3(1024), 0, s, part,
                     splits, count, out);
  return hipGetLastError();
}

hipError_t launch_channel_sum(const bf16_t* dYT, int N, int HW, int C, float* part, int splits,
                              float* db, hipStream_t s) {
  const int64_t total = (int64_t)N * HW;
  const int64_t chunk = (total + splits - 1) / splits;
  hipLaunchKernelGGL(k_channel_partial, dim3(C, splits), dim3(256), 0, s, dYT, N, HW, C, chunk,
                     part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_sum_splits(part, splits, C, db, s);
}

hipError_t launch_maxpool_fwd(const bf16_t* X, int N, int H, int W, int C, bf16_t* Y,
                              hipStream_t s) {
  if (C % 8 != 0 || (int64_t)N * H * W * C / 8 >= (1ll << 30)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_maxpool_fwd, grid1d((int64_t)N * (H / 2) * (W / 2) * C / 8), dim3(256), 0,
                     s, X, N, H, W, C, Y);
  return hipGetLastError();
}

hipError_t launch_maxpool_bwd(const bf16_t* X, const bf16_t* dY, int N, int H, int W, int C,
                              bf16_t* dX, hipStream_t s) {
  if (C % 8 != 0 || (int64_t)N * H * W * C / 8 >= (1ll << 30)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_maxpool_bwd, grid1d((int64_t)N * (H / 2) * (W / 2) * C / 8), dim3(256), 0,
                     s, X, dY, N, H, W, C, dX);
  return hipGetLastError();
}

hipError_t launch_upcat_fwd(const bf16_t* U, int Cu, const bf16_t* S, int Cs, int N, int H,
                            int W, bf16_t* cat, hipStream_t s) {
  if (Cu % 8 != 0 || Cs % 8 != 0 || (int64_t)N * H * W * (Cu + Cs) / 8 >= (1ll << 30))
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_upcat_fwd, grid1d((int64_t)N * H * W * (Cu + Cs) / 8), dim3(256), 0, s, U, Cu,
                     S, Cs, N, H, W, cat);
  return hipGetLastError();
}

hipError_t launch_upcat_bwd(const bf16_t* dcat, int Cu, int Cs, int N, int H, int W, bf16_t* dU,
                            bf16_t* dS, int accumulate, hipStream_t s) {
  if (Cu % 8 != 0 || Cs % 8 != 0 || (int64_t)N * H * W * (Cu + Cs) / 8 >= (1ll << 30))
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_upcat_bwd_u, grid1d((int64_t)N * (H / 2) * (W / 2) * Cu / 8), dim3(256), 0,
                     s, dcat, Cu, Cs, N, H, W, dU);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_upcat_bwd_s, grid1d((int64_t)N * H * W * Cs / 8), dim3(256), 0, s, dcat, Cu,
                     Cs, (int64_t)N * H * W, dS, accumulate);
  return hipGetLastError();
}

hipError_t launch_head_fwd(const bf16_t* Y, int64_t P, int C, const float* w, const float* b,
                           float* out, hipStream_t s) {
  switch (C) {
    case 16: hipLaunchKernelGGL(k_head_fwd<16>, grid1d(P), dim3(256), 0, s, Y, P, w, b, out); break;
    case 32: hipLaunchKernelGGL(k_head_fwd<32>, grid1d(P), dim3(256), 0, s, Y, P, w, b, out); break;
    case 64: hipLaunchKernelGGL(k_head_fwd<64>, grid1d(P), dim3(256), 0, s, Y, P, w, b, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_head_bwd(const bf16_t* Y, const float* dout, int64_t P, int C, const float* w,
                           bf16_t* dY, float* part, int nblk, hipStream_t s) {
  switch (C) {
    case 16: hipLaunchKernelGGL(k_head_bwd<16>, dim3(nblk), dim3(kHeadThreads), 0, s, Y, dout, P, w, dY, part); break;
    case 32: hipLaunchKernelGGL(k_head_bwd<32>, dim3(nblk), dim3(kHeadThreads), 0, s, Y, dout, P, w, dY, part); break;
    case 64: hipLaunchKernelGGL(k_head_bwd<64>, dim3(nblk), dim3(kHeadThreads), 0, s, Y, dout, P, w, dY, part); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_mse(const float* out, const float* target, int64_t P, float* dout,
                      double* part, int nblk, hipStream_t s) {
  hipLaunchKernelGGL(k_mse, dim3(nblk), dim3(kMseThreads), 0, s, out, target, P, dout, part);
  return hipGetLastError();
}

hipError_t launch_pack_input(const float* in, int64_t P, bf16_t* X8, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_input, grid1d(P), dim3(256), 0, s, in, P, X8);
  return hipGetLastError();
}

hipError_t launch_add_bf16(bf16_t* dS, const bf16_t* dX, int64_t count, hipStream_t s) {
  hipLaunchKernelGGL(k_add_bf16, grid1d(count), dim3(256), 0, s, dS, dX, count);
  return hipGetLastError();
}

hipError_t launch_adam(float* p, const float* g, float* m, float* v, int64_t count, float lr,
                       float b1, float b2, float eps, float c1, float c2, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, grid1d(count), dim3(256), 0, s, p, g, m, v, count, lr, b1, b2, eps,
                     c1, c2);
  return hipGetLastError();
}

hipError_t launch_to_bf16(const float* x, int64_t count, bf16_t* y, hipStream_t s) {
  hipLaunchKernelGGL(k_to_bf16, grid1d(count), dim3(256), 0, s, x, count, y);
  return hipGetLastError();
}

}  // namespace huvk
