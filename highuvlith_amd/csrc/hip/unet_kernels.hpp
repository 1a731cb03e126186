// Launchers of the U-Net kernels (unet_kernels.hip): BF16 storage, FP32
// accumulation on the matrix cores (v_mfma_f32_32x32x16_bf16).  Activations
// are NHWC; conv weights [Cout][3][3][Cin].  All pointers are device
// pointers; every launcher is asynchronous on `s`.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace huvk {

typedef uint16_t bf16_t;  // raw BF16 bits

// Y = act(conv3x3(X) + bias), pad 1, stride 1.  Cin % 8 == 0.  relu: 0/1.
// bias may be NULL.  Y has Cout channels (any Cout >= 1).
hipError_t launch_conv3x3_fwd(const bf16_t* X, int N, int H, int W, int Cin, const bf16_t* Wt,
                              const float* bias, int Cout, int relu, bf16_t* Y, hipStream_t s);
// Wr[ci][t][co] = W[co][8 - t][ci] (the kernel of the data gradient).
hipError_t launch_conv_weight_rot(const bf16_t* Wt, int Cout, int Cin, bf16_t* Wr, hipStream_t s);
// XT[n][c][y][x] = X[n][y][x][c] (H W a multiple of 64, C of 8; likewise below).
hipError_t launch_nhwc_to_nchw(const bf16_t* X, int N, int H, int W, int C, bf16_t* XT,
                               hipStream_t s);
// dY *= (Y > 0) in place (NHWC) and dYT = the masked dY as NCHW.  Y may be
// NULL (no ReLU).
hipError_t launch_relu_mask_t(bf16_t* dY, const bf16_t* Y, int N, int H, int W, int C,
                              bf16_t* dYT, hipStream_t s);
// Split-K weight gradient partials: part[split][co][t][ci] (FP32), W % 16 == 0,
// Cin % 8 == 0; rows (n, y) are split into `splits` contiguous ranges.
hipError_t launch_conv3x3_wgrad(const bf16_t* dYT, const bf16_t* XT, int N, int H, int W,
                                int Cin, int Cout, int splits, float* part, hipStream_t s);
// Weight gradient partials part[split][co][t][ci] and, when bpart is set, bias
// gradient partials bpart[split][co] straight from the NHWC tensors dY (ReLU
// already applied) and X; W % 64 == 0, Cin and Cout multiples of 8; rows
// (n, y) are split into `splits` contiguous ranges.
hipError_t launch_conv3x3_wgrad_nhwc(const bf16_t* dY, const bf16_t* X, int N, int H, int W,
                                     int Cin, int Cout, int splits, float* part, float* bpart,
                                     hipStream_t s);
// dY *= (Y > 0) in place (NHWC, C % 8 == 0).
hipError_t launch_relu_mask(bf16_t* dY, const bf16_t* Y, int N, int H, int W, int C,
                            hipStream_t s);
// out[i] = sum_s part[s][i] (fixed order), i < count.
hipError_t launch_sum_splits(const float* part, int splits, int64_t count, float* out,
                             hipStream_t s);
// db[c] = sum over n, pixels of dYT[n][c][.]  (NCHW): `splits` chunk
// partials in part (splits x C floats), then summed in a fixed order.
hipError_t launch_channel_sum(const bf16_t* dYT, int N, int HW, int C, float* part, int splits,
                              float* db, hipStream_t s);
// 2 x 2 max pooling (H, W even) and its gradient (the first maximum takes it).
hipError_t launch_maxpool_fwd(const bf16_t* X, int N, int H, int W, int C, bf16_t* Y,
                              hipStream_t s);
hipError_t launch_maxpool_bwd(const bf16_t* X, const bf16_t* dY, int N, int H, int W, int C,
                              bf16_t* dX, hipStream_t s);
// cat[n][y][x] = [nearest-2x(U)[n][y][x][0..Cu), S[n][y][x][0..Cs)] at (H, W);
// U is (H/2, W/2).  Backward: dU = sum of the 4 children, dS = dcat's tail
// (added to dS when accumulate != 0).
hipError_t launch_upcat_fwd(const bf16_t* U, int Cu, const bf16_t* S, int Cs, int N, int H,
                            int W, bf16_t* cat, hipStream_t s);
hipError_t launch_upcat_bwd(const bf16_t* dcat, int Cu, int Cs, int N, int H, int W, bf16_t* dU,
                            bf16_t* dS, int accumulate, hipStream_t s);
// out[p] = sum_c Y[p][c] w[c] + b (FP32); C in {16, 32, 64}.
hipError_t launch_head_fwd(const bf16_t* Y, int64_t P, int C, const float* w, const float* b,
                           float* out, hipStream_t s);
// dY[p][c] = dout[p] w[c]; partial sums for dw (C per block) and db.
hipError_t launch_head_bwd(const bf16_t* Y, const float* dout, int64_t P, int C, const float* w,
                           bf16_t* dY, float* part, int nblk, hipStream_t s);
// loss partials (FP64) and dout = 2 (out - target) / P.
hipError_t launch_mse(const float* out, const float* target, int64_t P, float* dout,
                      double* part, int nblk, hipStream_t s);
// X8[n][y][x][0..8) = (in[n][y][x], 0, ..., 0) as BF16.
hipError_t launch_pack_input(const float* in, int64_t P, bf16_t* X8, hipStream_t s);
// dS += dX (BF16, element-wise).
hipError_t launch_add_bf16(bf16_t* dS, const bf16_t* dX, int64_t count, hipStream_t s);
// Adam on the FP32 master parameters; BF16 copies refreshed for conv weights.
hipError_t launch_adam(float* p, const float* g, float* m, float* v, int64_t count, float lr,
                       float b1, float b2, float eps, float c1, float c2, hipStream_t s);
hipError_t launch_to_bf16(const float* x, int64_t count, bf16_t* y, hipStream_t s);

}  // namespace huvk
