// Team-level FFTs of length N = N1 * N2 (N1 <= N2 <= 64) by the four-step
// method: N1 register FFTs of length N2 over stride-N1 subsequences, a
// twiddle by W_N^{n1 k2}, an LDS transpose, then N2 register FFTs of length
// N1.  A team is N2 consecutive threads of a work-group; every thread of the
// work-group must call these together (they contain work-group barriers).
#pragma once

#include "fft_core.hpp"

namespace huvk {

// W_N^{k} = e^{-2 pi i k / N} table lookup, conjugated for the inverse.
template <int SIGN>
__device__ __forceinline__ c32 tw_lookup(const float2* __restrict__ tw, int k) {
  const float2 w = tw[k];
  return SIGN < 0 ? c32{w.x, w.y} : c32{w.x, -w.y};
}

template <int SIGN, int N>
__device__ __forceinline__ void fft_small(c32* v) {
  fft_reg<SIGN, N>(v);
}

// Threads per team for a four-step FFT of N1 x N2.
template <int N1, int N2>
constexpr int team_size() { return N1 == 1 ? 1 : N2; }

// Four-step FFT.  load(i) returns input element i (0 <= i < N); store(k, v)
// receives output element k.  scratch: per-team LDS area of (N1 + 1) * N2
// complex values.  tw: W_N table of N entries.  t: thread index in team.
template <int SIGN, int N1, int N2, class Load, class Store>
__device__ __forceinline__ void fft_four_step(int t, c32* scratch, const float2* __restrict__ tw,
                                              Load load, Store store) {
  constexpr int N = N1 * N2;
  if constexpr (N1 == 1) {
    c32 v[N2];
#pragma unroll
    for (int i = 0; i < N2; ++i) v[i] = load(i);  // only thread 0 of a 1-thread team
    fft_small<SIGN, N2>(v);
#pragma unroll
    for (int k = 0; k < N2; ++k) store(k, v[k]);
    (void)t; (void)scratch; (void)tw;
  } else {
    if (t < N1) {
      c32 v[N2];
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) v[n2] = load(t + N1 * n2);
      fft_small<SIGN, N2>(v);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        if (k2 > 0) v[k2] = cmul(v[k2], tw_lookup<SIGN>(tw, (t * k2) % N));
        scratch[t + (N1 + 1) * k2] = v[k2];
      }
    }
    __syncthreads();
    if (t < N2) {
      c32 u[N1];
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) u[n1] = scratch[n1 + (N1 + 1) * t];
      fft_small<SIGN, N1>(u);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) store(t + N2 * k1, u[k1]);
    }
    __syncthreads();
  }
}

}  // namespace huvk
