// In-LDS radix-4 Stockham FFT of 16 lines at a time, shared by the pruned real FFTs of the NUFFT
// fine grid (fft_pruned.hip) and the pruned complex Toeplitz operator of the mean CG
// (toeplitz.hip).
#pragma once
#include "common.h"

namespace cov3d {

constexpr int kFftThreads = 256;
constexpr int kFftLines = 16;
constexpr int kColGroup = 8;  // strided-pass loads in flight per thread

// In-place unnormalised FFT (sign -1 or +1) of kFftLines lines of length NF in LDS; line l
// starts at buf[l * LS].  tw[k] = exp(-2 pi i k / NF).  Stockham radix 4 (and 2): stage with
// stride Ns reads x[j + r NF/R], applies exp(sign 2 pi i r (j mod Ns) / (Ns R)), takes the R-point
// DFT and writes y[(j / Ns) Ns R + j mod Ns + r Ns].
template <int NF, int LS, int R>
__device__ __forceinline__ void fft_stage(float2* buf, const float2* tw, float sign, int Ns) {
  constexpr int Q = NF / R;
  constexpr int PER = kFftLines * Q / kFftThreads;
  static_assert(PER >= 1 && kFftLines * Q % kFftThreads == 0, "stage does not tile the block");
  const int tid = threadIdx.x;
  float2 v[PER][R];
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    const int g = tid + b * kFftThreads;
    const int l = g / Q, j = g % Q;
    const float2* x = buf + l * LS;
#pragma unroll
    for (int r = 0; r < R; ++r) v[b][r] = x[j + r * Q];
    const int jm = j & (Ns - 1);
    const int step = NF / (Ns * R);
#pragma unroll
    for (int r = 1; r < R; ++r) {
      const float2 w = tw[(r * jm * step) & (NF - 1)];
      const float wy = sign < 0.f ? w.y : -w.y;
      const float2 a = v[b][r];
      v[b][r] = make_float2(a.x * w.x - a.y * wy, a.x * wy + a.y * w.x);
    }
    if constexpr (R == 4) {
      // omega = exp(sign 2 pi i / 4) = sign * i; omega * (a + ib) = sign * (-b + ia)
      const float2 s02 = make_float2(v[b][0].x + v[b][2].x, v[b][0].y + v[b][2].y);
      const float2 d02 = make_float2(v[b][0].x - v[b][2].x, v[b][0].y - v[b][2].y);
      const float2 s13 = make_float2(v[b][1].x + v[b][3].x, v[b][1].y + v[b][3].y);
      const float2 d13 = make_float2(v[b][1].x - v[b][3].x, v[b][1].y - v[b][3].y);
      const float2 od13 = make_float2(-sign * d13.y, sign * d13.x);
      v[b][0] = make_float2(s02.x + s13.x, s02.y + s13.y);
      v[b][1] = make_float2(d02.x + od13.x, d02.y + od13.y);
      v[b][2] = make_float2(s02.x - s13.x, s02.y - s13.y);
      v[b][3] = make_float2(d02.x - od13.x, d02.y - od13.y);
    } else {
      const float2 a = v[b][0], c = v[b][1];
      v[b][0] = make_float2(a.x + c.x, a.y + c.y);
      v[b][1] = make_float2(a.x - c.x, a.y - c.y);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    const int g = tid + b * kFftThreads;
    const int l = g / Q, j = g % Q;
    const int jm = j & (Ns - 1);
    float2* y = buf + l * LS + (j / Ns) * Ns * R + jm;
#pragma unroll
    for (int r = 0; r < R; ++r) y[r * Ns] = v[b][r];
  }
  __syncthreads();
}

template <int NF, int LS>
__device__ void lds_fft(float2* buf, const float2* tw, float sign) {
  int Ns = 1;
  for (; Ns * 4 <= NF; Ns *= 4) fft_stage<NF, LS, 4>(buf, tw, sign, Ns);
  if (Ns < NF) fft_stage<NF, LS, 2>(buf, tw, sign, Ns);
}

// Fine-grid index of mode i (centred modes, negative frequencies wrapped to the end).
__device__ __forceinline__ int wrap_mode(int i, int N, int NF) {
  const int n = i - N / 2;
  return n >= 0 ? n : n + NF;
}

template <int NF>
__device__ __forceinline__ void load_twiddles(float2* s_tw, const float2* __restrict__ tw) {
  for (int k = threadIdx.x; k < NF; k += kFftThreads) s_tw[k] = tw[k];
}

// Twiddles exp(-2 pi i k / NF) evaluated in the kernel (no table in device memory).
template <int NF>
__device__ __forceinline__ void make_twiddles(float2* s_tw) {
  for (int k = threadIdx.x; k < NF; k += kFftThreads) {
    float sn, cs;
    sincospif(2.0f * (float)k / (float)NF, &sn, &cs);
    s_tw[k] = make_float2(cs, -sn);
  }
}

}  // namespace cov3d
