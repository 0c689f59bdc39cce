// Pruned 3-D real FFTs of the NUFFT fine grid (FP32, cubic grids of NF = 64,
// 128, 256 or 512 cells per axis, N <= NF/2 + 1 modes per axis).
//
// Type 2 turns the real modes into the Hermitian half of the fine-grid
// spectrum.  The zero-padded grid holds N of NF values on each axis, so the
// transform runs one axis at a time on the non-zero lines only, with the
// deconvolution and zero padding done while a line is loaded:
//   A: z lines of the N^2 mode rows -> (t, i0, i1, k2 <= NF/2)        sign -1
//   B: y lines (N of NF non-zero)   -> (t, i0, k1, k2)               sign -1
//   C: x lines (N of NF non-zero)   -> fine (t, k0, k1, k2)          sign -1
// which is the forward real-to-complex transform of the padded grid (the
// output of rocFFT's R2C of the same grid, to rounding).
//
// Type 1 (real output) folds the spread grid G into its Hermitian part
// H[k] = (G[k] + conj(G[-k])) / 2 while loading the first pass, and keeps only
// the N cropped outputs of every line:
//   X: x lines of H (k2 <= NF/2)    -> (t, i0, k1, k2)               sign +1
//   Y: y lines                      -> (t, i0, i1, k2)               sign +1
//   Z: Hermitian z lines -> real, deconvolved modes f (t, i0, i1, i2)
// which equals rocFFT's complex-to-real transform of H followed by the crop.
//
// The same passes give the centred 3-D transform of a zero-padded real volume,
// fftshift(fftn(ifftshift(pad_NF(v / gc)))) for a separable correction gc (the
// Fourier-domain eigenvolumes of the discretised projection): the volume's
// samples are the modes (centred index i - N/2, wrapped), 1/gc the per-axis
// weights, and the last pass (mode 4) writes the centred full grid, the
// mirror -k of every k2 < NF/2 column as the conjugate.  The adjoint (real
// part of the centred inverse transform, cropped and weighted) is type 1 with
// a first pass (mode 5) that reads the centred full grid.
// Every line is transformed in LDS by a radix-4 Stockham FFT (one radix-2
// stage when log2 NF is odd), 16 lines per workgroup.  The strided passes
// take 16 consecutive k2 columns per workgroup, so each row of a tile is one
// 128-byte segment in HBM.
#include <cmath>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "fft_lds.h"

namespace cov3d {

// ---- type 2, pass A: real mode rows (t, i0, i1, :) -> z spectrum (t, i0, i1, k2 <= NF/2).
// Lines g = (t * N + i0) * N + i1; 16 consecutive lines per workgroup.
template <int NF>
__global__ void __launch_bounds__(kFftThreads)
    k_pfft2_z(const float* __restrict__ f, float2* __restrict__ A, const float2* __restrict__ tw,
              const float* __restrict__ ip0, const float* __restrict__ ip1, const float* __restrict__ ip2,
              int64_t nlines, int64_t nlive, int N) {
  constexpr int LS = NF + 1;
  constexpr int H = NF / 2 + 1;
  __shared__ float2 s_buf[kFftLines * LS];
  __shared__ float2 s_tw[NF];
  load_twiddles<NF>(s_tw, tw);
  const int64_t g0 = (int64_t)blockIdx.x * kFftLines;
  for (int e = threadIdx.x; e < kFftLines * NF; e += kFftThreads) {
    const int l = e / NF, m = e % NF;
    s_buf[l * LS + m] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  // all of a thread's (at most NE) mode loads are issued before any is stored
  constexpr int NE = (kFftLines * H + kFftThreads - 1) / kFftThreads;
  float vin[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    const int e = threadIdx.x + k * kFftThreads;
    vin[k] = 0.f;
    if (e < kFftLines * N) {
      const int l = e / N, i2 = e % N;
      const int64_t g = g0 + l;
      if (g < nlines && g < nlive) vin[k] = f[g * N + i2];
    }
  }
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    const int e = threadIdx.x + k * kFftThreads;
    if (e < kFftLines * N) {
      const int l = e / N, i2 = e % N;
      const int64_t g = g0 + l;
      if (g < nlines) {
        const int i1 = (int)(g % N), i0 = (int)((g / N) % N);
        s_buf[l * LS + wrap_mode(i2, N, NF)] = make_float2(vin[k] * (ip0[i0] * ip1[i1] * ip2[i2]), 0.f);
      }
    }
  }
  __syncthreads();
  lds_fft<NF, LS>(s_buf, s_tw, -1.f);
  for (int e = threadIdx.x; e < kFftLines * H; e += kFftThreads) {
    const int l = e / H, k = e % H;
    const int64_t g = g0 + l;
    if (g < nlines) A[g * H + k] = s_buf[l * LS + k];
  }
}

// ---- strided passes: a workgroup takes 16 consecutive k2 columns [c0, c0 + 16) of one
// (outer, mid) position.  Input lines run over `nin` rows with row stride `sin` (complex
// elements); row i of the line goes to position pos(i).  Output rows o < nout come from line
// position opos(o) and go to row stride `sout`.
//   type 2, pass B: outer t*N+i0, input rows i1 < N at wrap(i1), output all NF rows
//   type 2, pass C: outer t, mid k1; input rows i0 < N at wrap(i0), output all NF rows
//   type 1, pass X: input all NF rows (folded), output rows i0 < N from wrap(i0)
//   type 1, pass Y: input all NF rows, output rows i1 < N from wrap(i1)
// MODE 0: pass B, 1: pass C, 2: pass X (with the Hermitian fold), 3: pass Y; 4: pass C writing the
// centred full grid, 5: pass X reading the centred full grid.
template <int NF, int MODE>
__global__ void __launch_bounds__(kFftThreads)
    k_pfft_cols(const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw,
                int ntrans, int N, int conj_out) {
  constexpr int LS = NF + 1;
  constexpr int H = NF / 2 + 1;
  constexpr int NCT = (H + kFftLines - 1) / kFftLines;  // column tiles
  __shared__ float2 s_buf[kFftLines * LS];
  __shared__ float2 s_tw[NF];
  load_twiddles<NF>(s_tw, tw);
  // the NCT column tiles of a row position (its 128-byte pieces, which straddle cache lines since
  // rows are H complex long) and neighbouring positions run on one XCD and meet in its L2
  const int64_t lb = xcd_contiguous(blockIdx.x, gridDim.x);
  const int ct = (int)(lb % NCT);
  const int64_t pos = lb / NCT;  // (outer, mid) position
  const int c0 = ct * kFftLines;
  const int ncol = min(kFftLines, H - c0);
  const int col = threadIdx.x % kFftLines, rsub = threadIdx.x / kFftLines;  // 16 columns x 16 rows
  constexpr float sign = (MODE <= 1 || MODE == 4) ? -1.f : 1.f;
  const bool in_pruned = MODE <= 1 || MODE == 4;  // input has N rows at wrapped positions, the rest zero
  constexpr int hN = NF / 2;
  constexpr int msk = NF - 1;
  // element (row, col) addresses; rows are complex elements apart by the row stride
  int64_t in_base, in_stride, out_base, out_stride;
  int nin = in_pruned ? N : NF;
  int nout = in_pruned ? NF : N;
  if constexpr (MODE == 0) {  // pos = t * N + i0; in A (t, i0, i1, k2), out B (t, i0, k1, k2)
    in_base = pos * N * H;
    in_stride = H;
    out_base = pos * NF * H;
    out_stride = H;
  } else if constexpr (MODE == 1 || MODE == 4) {
    // pos = t * NF + k1; in B (t, i0, k1, k2), out fine (t, k0, k1, k2) (mode 4: out unused)
    const int64_t t = pos / NF, k1 = pos % NF;
    in_base = (t * N * NF + k1) * H;
    in_stride = (int64_t)NF * H;
    out_base = (t * NF * NF + k1) * H;
    out_stride = (int64_t)NF * H;
  } else if constexpr (MODE == 2 || MODE == 5) {
    // pos = t * NF + k1; in G (t, k0, k1, k2 < NF) (mode 5: in unused), out P (t, i0, k1, k2)
    const int64_t t = pos / NF, k1 = pos % NF;
    in_base = (t * NF * NF + k1) * NF;
    in_stride = (int64_t)NF * NF;
    out_base = (t * N * NF + k1) * H;
    out_stride = (int64_t)NF * H;
  } else {  // pos = t * N + i0; in P (t, i0, k1, k2), out Q (t, i0, i1, k2)
    in_base = pos * NF * H;
    in_stride = H;
    out_base = pos * N * H;
    out_stride = H;
  }
  if (in_pruned) {
    for (int e = threadIdx.x; e < kFftLines * NF; e += kFftThreads) {
      const int l = e / NF, m = e % NF;
      s_buf[l * LS + m] = make_float2(0.f, 0.f);
    }
    __syncthreads();
  }
  if (col < ncol) {
    // rows rsub, rsub + 16, ...: GB rows' loads are issued before any is stored.  Measured with
    // the XCD-contiguous dealing: the passes whose rows are H apart (B, Y) take kColGroup; pass C
    // (and its centred form, mode 4), rows NF H apart, takes 4 at NF = 256 (0.75 -> 0.64 ms) and 8
    // at NF = 512 (8.78 -> 6.06 ms, where 65 KB of LDS leave two workgroups per CU); pass X keeps
    // one row at NF = 256 (4: 0.877 -> 0.889 ms) and takes 8 at NF = 512 (7.45 -> 7.26 ms).
    constexpr int RS = kFftThreads / kFftLines;
    constexpr int NIT = NF / RS;
    constexpr int GB = (MODE == 0 || MODE == 3) ? kColGroup
                       : (MODE == 1 || MODE == 4) ? (NF >= 512 ? 8 : 4)
                       : (MODE == 2 && NF >= 512) ? 8
                                                   : 1;
    for (int g0 = 0; g0 < NIT; g0 += GB) {
      float2 va[GB], vb[GB];
#pragma unroll
      for (int k = 0; k < GB; ++k) {
        const int i = rsub + (g0 + k) * RS;
        if (g0 + k < NIT && i < nin) {
          if constexpr (MODE == 2) {
            // H[k] = (G[k] + conj(G[-k])) / 2 at k = (i, k1, c0 + col), conjugated when conj_out
            const int64_t t = pos / NF, k1 = pos % NF;
            const int k2 = c0 + col;
            const int n0 = i ? NF - i : 0, n1 = k1 ? NF - (int)k1 : 0, n2 = k2 ? NF - k2 : 0;
            va[k] = in[in_base + i * in_stride + k2];
            vb[k] = in[(t * NF * NF + ((int64_t)n0 * NF + n1)) * NF + n2];
          } else if constexpr (MODE == 5) {
            // the same fold on the centred grid: k sits at (k + NF/2) mod NF on every axis
            const int64_t t = pos / NF;
            const int k1 = (int)(pos % NF);
            const int k2 = c0 + col;
            const int64_t tb = t * NF * NF * NF;
            va[k] = in[tb + ((int64_t)((i + hN) & msk) * NF + ((k1 + hN) & msk)) * NF + ((k2 + hN) & msk)];
            vb[k] = in[tb + ((int64_t)((hN - i) & msk) * NF + ((hN - k1) & msk)) * NF + ((hN - k2) & msk)];
          } else {
            va[k] = in[in_base + i * in_stride + c0 + col];
          }
        }
      }
#pragma unroll
      for (int k = 0; k < GB; ++k) {
        const int i = rsub + (g0 + k) * RS;
        if (g0 + k < NIT && i < nin) {
          const int m = in_pruned ? wrap_mode(i, N, NF) : i;
          float2 v = va[k];
          if constexpr (MODE == 2 || MODE == 5) {
            const float im = 0.5f * (v.y - vb[k].y);
            v = make_float2(0.5f * (v.x + vb[k].x), conj_out ? -im : im);
          }
          s_buf[col * LS + m] = v;
        }
      }
    }
  }
  __syncthreads();
  lds_fft<NF, LS>(s_buf, s_tw, sign);
  if constexpr (MODE == 4) {
    // centred full grid: k = (o, k1, k2) at (k + NF/2) mod NF; for 0 < k2 < NF/2 the mirror -k
    // (never computed by any workgroup) is written as the conjugate
    if (col < ncol) {
      const int64_t t = pos / NF;
      const int k1 = (int)(pos % NF);
      const int k2 = c0 + col;
      const int64_t tb = t * NF * NF * NF;
      const int k1c = (k1 + hN) & msk, k1m = (hN - k1) & msk;
      const bool mirror = k2 > 0 && k2 < hN;
      for (int o = rsub; o < NF; o += kFftThreads / kFftLines) {
        const float2 v = s_buf[col * LS + o];
        out[tb + ((int64_t)((o + hN) & msk) * NF + k1c) * NF + ((k2 + hN) & msk)] = v;
        if (mirror) out[tb + ((int64_t)((hN - o) & msk) * NF + k1m) * NF + (hN - k2)] = make_float2(v.x, -v.y);
      }
    }
  } else if (col < ncol) {
    for (int o = rsub; o < nout; o += kFftThreads / kFftLines) {
      const int m = in_pruned ? o : wrap_mode(o, N, NF);
      out[out_base + o * out_stride + c0 + col] = s_buf[col * LS + m];
    }
  }
}

// ---- type 1, pass Z: Hermitian z lines Q (t, i0, i1, k2 <= NF/2) -> real, cropped, deconvolved
// modes f (t, i0, i1, i2).  Lines g = (t * N + i0) * N + i1; 16 per workgroup.
template <int NF>
__global__ void __launch_bounds__(kFftThreads)
    k_pfft1_z(const float2* __restrict__ Q, float* __restrict__ f, const float2* __restrict__ tw,
              const float* __restrict__ ip0, const float* __restrict__ ip1, const float* __restrict__ ip2,
              int64_t nlines, int N) {
  constexpr int LS = NF + 1;
  constexpr int H = NF / 2 + 1;
  __shared__ float2 s_buf[kFftLines * LS];
  __shared__ float2 s_tw[NF];
  load_twiddles<NF>(s_tw, tw);
  const int64_t g0 = (int64_t)blockIdx.x * kFftLines;
  // all of a thread's (at most NE) loads are issued before any is stored
  constexpr int NE = (kFftLines * H + kFftThreads - 1) / kFftThreads;
  float2 vq[NE];
#pragma unroll
  for (int j = 0; j < NE; ++j) {
    const int e = threadIdx.x + j * kFftThreads;
    vq[j] = make_float2(0.f, 0.f);
    if (e < kFftLines * H && g0 + e / H < nlines) vq[j] = Q[(g0 + e / H) * H + e % H];
  }
#pragma unroll
  for (int j = 0; j < NE; ++j) {
    const int e = threadIdx.x + j * kFftThreads;
    if (e < kFftLines * H) {
      const int l = e / H, k = e % H;
      const float2 v = vq[j];
      s_buf[l * LS + k] = v;
      if (k > 0 && k < NF - k) s_buf[l * LS + NF - k] = make_float2(v.x, -v.y);
    }
  }
  __syncthreads();
  lds_fft<NF, LS>(s_buf, s_tw, 1.f);
  for (int e = threadIdx.x; e < kFftLines * N; e += kFftThreads) {
    const int l = e / N, i2 = e % N;
    const int64_t g = g0 + l;
    if (g < nlines) {
      const int i1 = (int)(g % N), i0 = (int)((g / N) % N);
      f[g * N + i2] = s_buf[l * LS + wrap_mode(i2, N, NF)].x * (ip0[i0] * ip1[i1] * ip2[i2]);
    }
  }
}

template <int NF>
static int pruned_type2_nf(const float* f, float2* fine, float2* work, const float2* tw, const float* ip0,
                           const float* ip1, const float* ip2, int ntrans, int nrun, int N, hipStream_t st) {
  constexpr int H = NF / 2 + 1;
  constexpr int NCT = (H + kFftLines - 1) / kFftLines;
  float2* A = work;                                   // (ntrans, N, N, H)
  float2* B = work + (int64_t)ntrans * N * N * H;     // (ntrans, N, NF, H)
  const int64_t nlines = (int64_t)ntrans * N * N;
  hipLaunchKernelGGL((k_pfft2_z<NF>), dim3((unsigned)ceil_div(nlines, kFftLines)), dim3(kFftThreads), 0, st, f,
                     A, tw, ip0, ip1, ip2, nlines, (int64_t)nrun * N * N, N);
  COV3D_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_pfft_cols<NF, 0>), dim3((unsigned)((int64_t)ntrans * N * NCT)), dim3(kFftThreads), 0, st,
                     A, B, tw, ntrans, N, 0);
  COV3D_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_pfft_cols<NF, 1>), dim3((unsigned)((int64_t)ntrans * NF * NCT)), dim3(kFftThreads), 0, st,
                     B, fine, tw, ntrans, N, 0);
  COV3D_LAUNCH_CHECK();
  return 0;
}

template <int NF>
static int pruned_type1_nf(const float2* G, float* f, float2* work, const float2* tw, const float* ip0,
                           const float* ip1, const float* ip2, int ntrans, int N, int conj_out, hipStream_t st) {
  constexpr int H = NF / 2 + 1;
  constexpr int NCT = (H + kFftLines - 1) / kFftLines;
  float2* P = work;                                   // (ntrans, N, NF, H)
  float2* Q = work + (int64_t)ntrans * N * NF * H;    // (ntrans, N, N, H)
  hipLaunchKernelGGL((k_pfft_cols<NF, 2>), dim3((unsigned)((int64_t)ntrans * NF * NCT)), dim3(kFftThreads), 0, st,
                     G, P, tw, ntrans, N, conj_out);
  COV3D_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_pfft_cols<NF, 3>), dim3((unsigned)((int64_t)ntrans * N * NCT)), dim3(kFftThreads), 0, st,
                     P, Q, tw, ntrans, N, 0);
  COV3D_LAUNCH_CHECK();
  const int64_t nlines = (int64_t)ntrans * N * N;
  hipLaunchKernelGGL((k_pfft1_z<NF>), dim3((unsigned)ceil_div(nlines, kFftLines)), dim3(kFftThreads), 0, st, Q,
                     f, tw, ip0, ip1, ip2, nlines, N);
  COV3D_LAUNCH_CHECK();
  return 0;
}

// Centred transform of zero-padded real volumes: passes A, B and C (mode 4).
template <int NF>
static int fft3c_pad_nf(const float* f, float2* out, float2* work, const float2* tw, const float* ip, int ntrans,
                        int N, hipStream_t st) {
  constexpr int H = NF / 2 + 1;
  constexpr int NCT = (H + kFftLines - 1) / kFftLines;
  float2* A = work;                                   // (ntrans, N, N, H)
  float2* B = work + (int64_t)ntrans * N * N * H;     // (ntrans, N, NF, H)
  const int64_t nlines = (int64_t)ntrans * N * N;
  hipLaunchKernelGGL((k_pfft2_z<NF>), dim3((unsigned)ceil_div(nlines, kFftLines)), dim3(kFftThreads), 0, st, f,
                     A, tw, ip, ip, ip, nlines, nlines, N);
  COV3D_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_pfft_cols<NF, 0>), dim3((unsigned)((int64_t)ntrans * N * NCT)), dim3(kFftThreads), 0, st,
                     A, B, tw, ntrans, N, 0);
  COV3D_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_pfft_cols<NF, 4>), dim3((unsigned)((int64_t)ntrans * NF * NCT)), dim3(kFftThreads), 0, st,
                     B, out, tw, ntrans, N, 0);
  COV3D_LAUNCH_CHECK();
  return 0;
}

// Its adjoint: passes X (mode 5), Y and Z.
template <int NF>
static int fft3c_pad_adjoint_nf(const float2* G, float* f, float2* work, const float2* tw, const float* ip,
                                int ntrans, int N, hipStream_t st) {
  constexpr int H = NF / 2 + 1;
  constexpr int NCT = (H + kFftLines - 1) / kFftLines;
  float2* P = work;                                   // (ntrans, N, NF, H)
  float2* Q = work + (int64_t)ntrans * N * NF * H;    // (ntrans, N, N, H)
  hipLaunchKernelGGL((k_pfft_cols<NF, 5>), dim3((unsigned)((int64_t)ntrans * NF * NCT)), dim3(kFftThreads), 0, st,
                     G, P, tw, ntrans, N, 0);
  COV3D_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_pfft_cols<NF, 3>), dim3((unsigned)((int64_t)ntrans * N * NCT)), dim3(kFftThreads), 0, st,
                     P, Q, tw, ntrans, N, 0);
  COV3D_LAUNCH_CHECK();
  const int64_t nlines = (int64_t)ntrans * N * N;
  hipLaunchKernelGGL((k_pfft1_z<NF>), dim3((unsigned)ceil_div(nlines, kFftLines)), dim3(kFftThreads), 0, st, Q,
                     f, tw, ip, ip, ip, nlines, N);
  COV3D_LAUNCH_CHECK();
  return 0;
}

bool pruned_fft_supported(int64_t NF, int64_t N) {
  return (NF == 64 || NF == 128 || NF == 256 || NF == 512) && N >= 1 && 2 * N <= NF + 2;
}

// Work area (complex values) both pipelines need.
int64_t pruned_fft_work(int64_t NF, int64_t N, int64_t ntrans) {
  const int64_t H = NF / 2 + 1;
  return ntrans * N * H * (N + NF);
}

int pruned_type2(const float* f, void* fine, void* work, const void* tw, const float* ip0, const float* ip1,
                 const float* ip2, int ntrans, int nrun, int NF, int N, hipStream_t st) {
  float2* fo = (float2*)fine;
  float2* w = (float2*)work;
  const float2* t = (const float2*)tw;
  switch (NF) {
    case 64: return pruned_type2_nf<64>(f, fo, w, t, ip0, ip1, ip2, ntrans, nrun, N, st);
    case 128: return pruned_type2_nf<128>(f, fo, w, t, ip0, ip1, ip2, ntrans, nrun, N, st);
    case 256: return pruned_type2_nf<256>(f, fo, w, t, ip0, ip1, ip2, ntrans, nrun, N, st);
    case 512: return pruned_type2_nf<512>(f, fo, w, t, ip0, ip1, ip2, ntrans, nrun, N, st);
    default: set_error("pruned FFT: unsupported grid %d", NF); return COV3D_ERR_ARG;
  }
}

int pruned_type1(const void* G, float* f, void* work, const void* tw, const float* ip0, const float* ip1,
                 const float* ip2, int ntrans, int NF, int N, int conj_out, hipStream_t st) {
  const float2* g = (const float2*)G;
  float2* w = (float2*)work;
  const float2* t = (const float2*)tw;
  switch (NF) {
    case 64: return pruned_type1_nf<64>(g, f, w, t, ip0, ip1, ip2, ntrans, N, conj_out, st);
    case 128: return pruned_type1_nf<128>(g, f, w, t, ip0, ip1, ip2, ntrans, N, conj_out, st);
    case 256: return pruned_type1_nf<256>(g, f, w, t, ip0, ip1, ip2, ntrans, N, conj_out, st);
    case 512: return pruned_type1_nf<512>(g, f, w, t, ip0, ip1, ip2, ntrans, N, conj_out, st);
    default: set_error("pruned FFT: unsupported grid %d", NF); return COV3D_ERR_ARG;
  }
}

// exp(-2 pi i k / NF) tables, one per (device, NF), made on first use and kept
static std::mutex g_tw_mu;
static std::map<std::pair<int, int64_t>, float2*> g_tw;

static int twiddles(int64_t NF, const float2** out) {
  int dev = 0;
  COV3D_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto it = g_tw.find({dev, NF});
  if (it != g_tw.end()) {
    *out = it->second;
    return 0;
  }
  std::vector<float2> h(NF);
  for (int64_t k = 0; k < NF; ++k)
    h[k] = make_float2((float)cos(2.0 * M_PI * (double)k / (double)NF), (float)-sin(2.0 * M_PI * (double)k / (double)NF));
  float2* d = nullptr;
  COV3D_HIP(hipMalloc(&d, NF * sizeof(float2)));
  if (hipMemcpy(d, h.data(), NF * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    set_error("twiddle table upload failed");
    return COV3D_ERR_HIP;
  }
  g_tw[{dev, NF}] = d;
  *out = d;
  return 0;
}

static int fft3c_pad_check(int64_t ntrans, int64_t L, int64_t nf, const void* a, const void* b, const void* work,
                           const float* ip) {
  COV3D_CHECK_ARG(a && b && work && ip, "null pointer");
  COV3D_CHECK_ARG(ntrans >= 1 && ntrans <= (1 << 20), "ntrans %lld out of range", (long long)ntrans);
  COV3D_CHECK_ARG(pruned_fft_supported(nf, L), "padded grid %lld for L=%lld: needs 64, 128, 256 or 512 cells and "
                  "2L <= grid + 2", (long long)nf, (long long)L);
  return 0;
}

}  // namespace cov3d

extern "C" int64_t cov3d_fft3c_pad_real_work(int64_t L, int64_t nf, int64_t ntrans) {
  return cov3d::pruned_fft_supported(nf, L) ? cov3d::pruned_fft_work(nf, L, ntrans) : -1;
}

extern "C" int cov3d_fft3c_pad_real(const float* v, void* out, void* work, const float* ip, int64_t ntrans,
                                    int64_t L, int64_t nf, void* stream) {
  using namespace cov3d;
  if (int rc = fft3c_pad_check(ntrans, L, nf, v, out, work, ip)) return rc;
  const float2* tw = nullptr;
  if (int rc = twiddles(nf, &tw)) return rc;
  hipStream_t st = (hipStream_t)stream;
  float2* o = (float2*)out;
  float2* w = (float2*)work;
  switch (nf) {
    case 64: return fft3c_pad_nf<64>(v, o, w, tw, ip, (int)ntrans, (int)L, st);
    case 128: return fft3c_pad_nf<128>(v, o, w, tw, ip, (int)ntrans, (int)L, st);
    case 256: return fft3c_pad_nf<256>(v, o, w, tw, ip, (int)ntrans, (int)L, st);
    default: return fft3c_pad_nf<512>(v, o, w, tw, ip, (int)ntrans, (int)L, st);
  }
}

extern "C" int cov3d_fft3c_pad_real_adjoint(const void* g, float* v, void* work, const float* ip, int64_t ntrans,
                                            int64_t L, int64_t nf, void* stream) {
  using namespace cov3d;
  if (int rc = fft3c_pad_check(ntrans, L, nf, g, v, work, ip)) return rc;
  const float2* tw = nullptr;
  if (int rc = twiddles(nf, &tw)) return rc;
  hipStream_t st = (hipStream_t)stream;
  const float2* gi = (const float2*)g;
  float2* w = (float2*)work;
  switch (nf) {
    case 64: return fft3c_pad_adjoint_nf<64>(gi, v, w, tw, ip, (int)ntrans, (int)L, st);
    case 128: return fft3c_pad_adjoint_nf<128>(gi, v, w, tw, ip, (int)ntrans, (int)L, st);
    case 256: return fft3c_pad_adjoint_nf<256>(gi, v, w, tw, ip, (int)ntrans, (int)L, st);
    default: return fft3c_pad_adjoint_nf<512>(gi, v, w, tw, ip, (int)ntrans, (int)L, st);
  }
}
