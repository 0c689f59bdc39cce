// 3-D type-1 / type-2 NUFFT for gfx950 with the exponential-of-semicircle
// (ES) kernel and oversampling factor 2.
//
// Replaces the finufft / cufinufft plans driven by cov3d/nufft_plan.py:263-392.
// Method (published FINUFFT / cuFINUFFT design, re-derived here):
//   type 2: deconvolve + zero-pad modes onto the (nf0, nf1, nf2) fine grid,
//           FFT, then interpolate every point with a W^3 ES stencil;
//   type 1: spread every point with the W^3 ES stencil, FFT, deconvolve+crop.
// Points are bin-sorted once per setpts (stable radix sort by 4^3 bin and axis-0
// stencil position), so every per-bin sum runs over the same points in the same order.
//   interpolation (k_interp_pipe, FP32; k_interp otherwise): a workgroup stages one
//     super-bin's fine-grid tile plus its W-wide halo in LDS, two transforms per
//     16-byte cell, and a pair of lanes per point (FP32; one thread per point
//     otherwise) reads its W^3 taps from LDS and writes the caller's (ntrans, M) layout; the pipelined kernel is persistent, takes its
//     items from per-XCD counters and loads the next tile while the current taps run;
//   spreading (k_spread_mfma, FP32; k_spread_gather for FP64): output-driven gather,
//     one workgroup per 4^3 bin walks the points of the 27 neighbouring bins and forms
//     each cell as a matrix product (weights x values) on the FP32 matrix cores, so each
//     cell is written once, with no atomics, and the result is deterministic;
//   FFTs: pruned real 3-D FFTs on 64/128/256/512-cell grids (fft_pruned.hip), rocFFT
//     for FP64 and other grids.
#include <math.h>
#include <rocfft/rocfft.h>

#include <mutex>
#include <type_traits>
#include <vector>

#include "common.h"

namespace cov3d {

// ----------------------------------------------------------------------------
// Parameters
// ----------------------------------------------------------------------------

// Kernel width for a requested tolerance at oversampling 2: w = ceil(log10(1/eps)) + 1.
static int es_width(double eps) {
  int w = (int)ceil(log10(1.0 / eps)) + 1;
  if (w < 2) w = 2;
  if (w > 16) w = 16;
  return w;
}

// ES shape parameter beta = c(w) * w at oversampling 2.
static double es_beta(int w) {
  double b = 2.30;
  if (w == 2) b = 2.20;
  else if (w == 3) b = 2.26;
  else if (w == 4) b = 2.38;
  return b * w;
}

// Smallest even integer >= n whose only prime factors are 2, 3, 5.
static int64_t next235even(int64_t n) {
  if (n <= 2) return 2;
  if (n % 2) n++;
  for (int64_t m = n;; m += 2) {
    int64_t k = m;
    while (k % 2 == 0) k /= 2;
    while (k % 3 == 0) k /= 3;
    while (k % 5 == 0) k /= 5;
    if (k == 1) return m;
  }
}

// Gauss-Legendre nodes/weights on [-1, 1] (Newton on P_n).
static void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& wts) {
  x.resize(n);
  wts.resize(n);
  for (int i = 0; i < n; ++i) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5));
    double pp = 0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1, p2 = 0;
      for (int j = 1; j <= n; ++j) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1);
      double dz = p1 / pp;
      z -= dz;
      if (fabs(dz) < 1e-16) break;
    }
    x[i] = z;
    wts[i] = 2.0 / ((1 - z * z) * pp * pp);
  }
}

static inline double es_eval_host(double t, int w, double beta) {
  double c = 4.0 / ((double)w * w);
  double s = 1.0 - c * t * t;
  if (s < 0) return 0;
  return exp(beta * (sqrt(s) - 1.0));
}

// Degree-8 polynomial form of the stencil weights.  With j0 = ceil(u - W/2) the first tap of a
// point at u, tap a's weight is phi(x + a - W/2) for x = j0 - u + W/2 in [0, 1).  In y = x - 1/2,
// tap a is P_a(y) = E_a(y^2) + y O_a(y^2), and phi being even gives tap W-1-a = E_a(y^2) - y O_a(y^2):
// one pair of short Horner chains per mirrored pair of taps, no exp or sqrt.  P_a interpolates phi
// at the 9 Chebyshev points of [-1/2, 1/2] (solved here in FP64); evaluated in FP32 it is within
// 1.2e-7 of phi for W = 7 (the FP32 floor; the native exp/sqrt path is at the same level).
struct EsPoly {
  float e[4][5];  // E_a: coefficients of y^0, y^2, ..., y^8
  float o[4][4];  // O_a: coefficients of y^1, y^3, ..., y^7
  int on;         // 1: the FP32 interpolation kernels use this form instead of exp/sqrt
};

static EsPoly es_poly(int w, double beta) {
  EsPoly P{};
  constexpr int n = 9;
  for (int a = 0; a < (w + 1) / 2 && a < 4; ++a) {
    double A[n][n + 1];
    for (int k = 0; k < n; ++k) {
      const double y = 0.5 * cos(M_PI * (2 * k + 1) / (2.0 * n));
      double pw = 1.0;
      for (int d = 0; d < n; ++d, pw *= y) A[k][d] = pw;
      A[k][n] = es_eval_host(y + 0.5 + a - 0.5 * w, w, beta);
    }
    for (int col = 0; col < n; ++col) {  // Gaussian elimination, partial pivoting
      int piv = col;
      for (int r = col + 1; r < n; ++r)
        if (fabs(A[r][col]) > fabs(A[piv][col])) piv = r;
      for (int d = 0; d <= n; ++d) std::swap(A[col][d], A[piv][d]);
      for (int r = 0; r < n; ++r) {
        if (r == col) continue;
        const double f = A[r][col] / A[col][col];
        for (int d = col; d <= n; ++d) A[r][d] -= f * A[col][d];
      }
    }
    for (int d = 0; d < n; ++d) {
      const float cf = (float)(A[d][n] / A[d][d]);
      if (d % 2 == 0) P.e[a][d / 2] = cf;
      else P.o[a][d / 2] = cf;
    }
  }
  return P;
}

// 1/p(n) for n in [-floor(N/2), ceil(N/2)-1], p(n) = int phi(t) cos(2 pi n t / nf) dt.
static std::vector<double> deconv_factors(int64_t N, int64_t nf, int w, double beta) {
  std::vector<double> gx, gw;
  gauss_legendre(128, gx, gw);
  std::vector<double> out(N);
  double half = 0.5 * w;
  for (int64_t i = 0; i < N; ++i) {
    double n = (double)(i - N / 2);
    double s = 0;
    for (size_t q = 0; q < gx.size(); ++q) {
      double t = half * gx[q];  // nodes on [-w/2, w/2]
      s += half * gw[q] * es_eval_host(t, w, beta) * cos(2.0 * M_PI * n * t / (double)nf);
    }
    out[i] = 1.0 / s;
  }
  return out;
}

// ----------------------------------------------------------------------------
// Device kernels
// ----------------------------------------------------------------------------

constexpr int kBlock = 256;
constexpr int kMaxSub = 1024;  // points per subproblem

// ----------------------------------------------------------------------------
// Optional HIP-event timing of kernel classes on the stream they run on.
// ----------------------------------------------------------------------------
enum TimedKernel {
  kTInterp = 0,
  kTSpread = 1,
  kTDeconvPad = 2,
  kTDeconvCrop = 3,
  kTFft = 4,
  kTSort = 5,
  kTPermute = 6,
  kTCount
};

struct TimingRegistry {
  bool enabled = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[kTCount];
};
static TimingRegistry g_timing;

struct ScopedTiming {
  int k;
  hipStream_t st;
  hipEvent_t a = nullptr, b = nullptr;
  ScopedTiming(int kind, hipStream_t s) : k(kind), st(s) {
    if (!g_timing.enabled) return;
    hipEventCreate(&a);
    hipEventCreate(&b);
    hipEventRecord(a, st);
  }
  ~ScopedTiming() {
    if (!a) return;
    hipEventRecord(b, st);
    g_timing.ev[k].push_back({a, b});
  }
};

template <typename T>
__device__ __forceinline__ T es_eval(T t, T c, T beta) {
  T s = (T)1 - c * t * t;
  s = s > (T)0 ? s : (T)0;
  return exp(beta * (sqrt(s) - (T)1));
}

// ES kernel in FP32 on the native v_sqrt_f32 / v_exp_f32 (about 1 ulp each):
// exp(beta (sqrt(s) - 1)) = 2^(beta log2(e) (sqrt(s) - 1)), beta_l2 = beta log2(e).
__device__ __forceinline__ float es_eval_native(float t, float c, float beta_l2) {
  const float s = fmaxf(1.f - c * t * t, 0.f);
  return __builtin_amdgcn_exp2f(beta_l2 * (__builtin_amdgcn_sqrtf(s) - 1.f));
}

__global__ void k_fill_int(int* p, int64_t n, int v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T>
__global__ void k_fill_zero(T* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = (T)0;
}


// Binning geometry.  Bins are 4^3 fine-grid cells; 2x2x2 bins form an 8^3
// super-bin (the interpolation tile).  Bins are numbered super-bin-major so
// the points of one super-bin are contiguous after sorting.  nf is a multiple
// of 8 on every axis, so bins tile the periodic grid exactly.
constexpr int kBinEdge = 4;
constexpr int kSuperEdge = 8;

struct BinGeom {
  int nf[3];
  int nb[3];   // bins per axis (nf / 4)
  int nsb[3];  // super-bins per axis (nf / 8)
  // spreading: for XCD x, xcd_tiles[x * n + j] is the tile ordinal (gather_bin's numbering) it
  // takes j-th, nearest the grid centre first; NULL keeps the ordinal order
  const int* xcd_tiles;
};

// Super-bin numbering.  When every axis holds a multiple of 4 super-bins they are numbered in
// 4x4x4 blocks (block-major, raster inside a block), otherwise in raster order.  Subproblems follow
// this order, and the interpolation deals runs of consecutive subproblems to one XCD, so a run is
// a compact block whose tiles share their halos in that XCD's L2 (a run of two raster rows shares
// them along two axes only).  Every per-point and per-bin sum is independent of the numbering.
__host__ __device__ __forceinline__ bool super_blocked(const BinGeom& g) {
  return (g.nsb[0] % 4) == 0 && (g.nsb[1] % 4) == 0 && (g.nsb[2] % 4) == 0;
}

__host__ __device__ __forceinline__ int super_index(const BinGeom& g, int s0, int s1, int s2) {
  if (!super_blocked(g)) return (s0 * g.nsb[1] + s1) * g.nsb[2] + s2;
  const int blk = ((s0 >> 2) * (g.nsb[1] >> 2) + (s1 >> 2)) * (g.nsb[2] >> 2) + (s2 >> 2);
  return blk * 64 + (((s0 & 3) << 4) | ((s1 & 3) << 2) | (s2 & 3));
}

__host__ __device__ __forceinline__ void super_coords(const BinGeom& g, int sb, int& s0, int& s1, int& s2) {
  if (!super_blocked(g)) {
    s2 = sb % g.nsb[2];
    s1 = (sb / g.nsb[2]) % g.nsb[1];
    s0 = sb / (g.nsb[2] * g.nsb[1]);
    return;
  }
  const int blk = sb >> 6, w = sb & 63;
  const int n1 = g.nsb[1] >> 2, n2 = g.nsb[2] >> 2;
  s0 = ((blk / (n1 * n2)) << 2) | (w >> 4);
  s1 = (((blk / n2) % n1) << 2) | ((w >> 2) & 3);
  s2 = ((blk % n2) << 2) | (w & 3);
}

__host__ __device__ __forceinline__ int bin_index(const BinGeom& g, int b0, int b1, int b2) {
  const int sb = super_index(g, b0 >> 1, b1 >> 1, b2 >> 1);
  return sb * 8 + ((b0 & 1) << 2) + ((b1 & 1) << 1) + (b2 & 1);
}

__device__ __forceinline__ void bin_coords(const BinGeom& g, int bin, int& b0, int& b1, int& b2) {
  int sb = bin >> 3, k = bin & 7;
  int s0, s1, s2;
  super_coords(g, sb, s0, s1, s2);
  b0 = 2 * s0 + ((k >> 2) & 1);
  b1 = 2 * s1 + ((k >> 1) & 1);
  b2 = 2 * s2 + (k & 1);
}

// Wrapped fine-grid coordinate X in [0, nf) (double), its bin, and the
// coordinate relative to the bin's super-bin origin, v = X - 8*floor(b/2).
template <typename T>
__device__ __forceinline__ int point_bin(const T* x, const T* y, const T* z, int64_t j,
                                         const BinGeom& g, double v[3]) {
  const double inv2pi = 0.15915494309189533577;
  double p[3] = {(double)x[j], (double)y[j], (double)z[j]};
  int b[3];
  for (int k = 0; k < 3; ++k) {
    double nf = (double)g.nf[k];
    double X = p[k] * inv2pi * nf;
    X -= nf * floor(X / nf);
    if (X >= nf) X -= nf;
    if (X < 0) X = 0;
    int bi = (int)floor(X) / kBinEdge;
    if (bi > g.nb[k] - 1) bi = g.nb[k] - 1;
    b[k] = bi;
    v[k] = X - (double)((bi >> 1) * kSuperEdge);
  }
  return bin_index(g, b[0], b[1], b[2]);
}

// Bin of every point, its sort key and its super-bin-relative coordinates, kept as a 4-wide
// record pv[4j..4j+2] for k_bin_place.  The key is bin * 8 + the point's first stencil cell
// along axis 0 relative to its bin (ceil(v0 - W/2) - bin start + 4, clamped to [0, 7]): within a
// bin the sorted points then run in groups of one axis-0 stencil position, so consecutive hits of
// the matrix-core spreading share their zero axis-0 weights and skip whole cell blocks together.
// Consecutive points (neighbouring pixels of one image) often share a bin: each run of equal bins
// among consecutive lanes adds its length with one atomic.
template <typename T>
__global__ void k_bin_count(const T* x, const T* y, const T* z, int64_t M, BinGeom g,
                            int* bin_count, int* pt_bin, T* __restrict__ pv, int W) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = j < M;
  int bin = -1;
  if (ok) {
    double v[3];
    bin = point_bin(x, y, z, j, g, v);
    int b0, b1, b2;
    bin_coords(g, bin, b0, b1, b2);
    const int xs = (int)ceil(v[0] - 0.5 * W) - kBinEdge * (b0 & 1) + 4;
    pt_bin[j] = bin * 8 + min(max(xs, 0), 7);
    pv[4 * j] = (T)v[0];
    pv[4 * j + 1] = (T)v[1];
    pv[4 * j + 2] = (T)v[2];
  }
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(bin, 1);
  const bool start = ok && (lane == 0 || prev != bin);
  const unsigned long long ends = __ballot(start || !ok);  // a lane past the end closes a run
  if (start) {
    const unsigned long long above = lane == 63 ? 0ull : ends >> (lane + 1);
    const int len = above ? __ffsll((long long)above) : 64 - lane;
    atomicAdd(&bin_count[bin], len);
  }
}

__global__ void k_sub_fill(const int* bin_count, const int* bin_start, const int* sub_start,
                           int64_t nsuper, int* sub_sb, int* sub_pt, int* sub_n) {
  int64_t sb = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sb >= nsuper) return;
  int c = 0;
  for (int k = 0; k < 8; ++k) c += bin_count[8 * sb + k];
  int s0 = sub_start[sb];
  int p0 = bin_start[8 * sb];
  for (int s = 0; s * kMaxSub < c; ++s) {
    sub_sb[s0 + s] = (int)sb;
    sub_pt[s0 + s] = p0 + s * kMaxSub;
    int rem = c - s * kMaxSub;
    sub_n[s0 + s] = rem < kMaxSub ? rem : kMaxSub;
  }
}

// Sorted position p holds point sidx[p] (stable radix sort by bin): its inverse and its
// super-bin-relative coordinates (from k_bin_count's records).
template <typename T>
__global__ void k_bin_place(const T* __restrict__ pv, int64_t M, const int* __restrict__ sidx, int* spos, T* sv0,
                            T* sv1, T* sv2) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= M) return;
  const int j = sidx[p];
  T v[4];
  if constexpr (sizeof(T) == 4) {
    const float4 r = reinterpret_cast<const float4*>(pv)[j];
    v[0] = r.x;
    v[1] = r.y;
    v[2] = r.z;
  } else {
    const double2 a = reinterpret_cast<const double2*>(pv)[2 * (int64_t)j];
    const double2 b = reinterpret_cast<const double2*>(pv)[2 * (int64_t)j + 1];
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
  }
  spos[j] = (int)p;
  sv0[p] = v[0];
  sv1[p] = v[1];
  sv2[p] = v[2];
}

hipError_t sort_points_by_bin(void* tmp, size_t& tmp_bytes, const unsigned* keys_in, unsigned* keys_out,
                              int* idx_out, int n, int nbits, hipStream_t st);
hipError_t exclusive_scan_ints(void* tmp, size_t& tmp_bytes, const int* in, int* out, int n, hipStream_t st);

// Subproblems of each super-bin (ceil(points / kMaxSub)), scanned into sub_start afterwards.
__global__ void k_sub_count(const int* __restrict__ bin_count, int64_t nsuper, int* __restrict__ sub_cnt) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsuper) return;
  int c = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) c += bin_count[8 * s + k];
  sub_cnt[s] = (c + kMaxSub - 1) / kMaxSub;
}

__global__ void k_sub_total(const int* __restrict__ sub_cnt, const int* __restrict__ sub_start, int64_t nsuper,
                            int* __restrict__ nsub) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *nsub = nsuper > 0 ? sub_start[nsuper - 1] + sub_cnt[nsuper - 1] : 0;
}

// fft_pruned.hip
bool pruned_fft_supported(int64_t NF, int64_t N);
int64_t pruned_fft_work(int64_t NF, int64_t N, int64_t ntrans);
int pruned_type2(const float* f, void* fine, void* work, const void* tw, const float* ip0, const float* ip1,
                 const float* ip2, int ntrans, int nrun, int NF, int N, hipStream_t st);
int pruned_type1(const void* G, float* f, void* work, const void* tw, const float* ip0, const float* ip1,
                 const float* ip2, int ntrans, int NF, int N, int conj_out, hipStream_t st);

// Modes -> zero-padded, deconvolved fine grid (one thread per fine-grid cell).
template <typename T, bool REAL_IN>
__global__ void k_deconv_pad(const T* __restrict__ f, Cplx<T>* __restrict__ fine, int64_t ntrans,
                             int64_t nrun, int N0, int N1, int N2, int nf0, int nf1, int nf2,
                             const T* __restrict__ ip0, const T* __restrict__ ip1,
                             const T* __restrict__ ip2) {
  const int64_t nfine = (int64_t)nf0 * nf1 * nf2;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nfine * ntrans) return;
  int64_t t = idx / nfine;
  int64_t r = idx - t * nfine;
  int m2 = (int)(r % nf2);
  int64_t r2 = r / nf2;
  int m1 = (int)(r2 % nf1);
  int m0 = (int)(r2 / nf1);
  // fine index m <-> mode n = m (m < ceil(N/2)) or m - nf (m >= nf - floor(N/2))
  auto mode_index = [](int m, int N, int nf) -> int {
    int hp = (N + 1) / 2, hn = N / 2;
    if (m < hp) return m + hn;
    if (m >= nf - hn) return m - nf + hn;
    return -1;
  };
  int i0 = mode_index(m0, N0, nf0), i1 = mode_index(m1, N1, nf1), i2 = mode_index(m2, N2, nf2);
  Cplx<T> v = {(T)0, (T)0};
  if (t < nrun && i0 >= 0 && i1 >= 0 && i2 >= 0) {
    T s = ip0[i0] * ip1[i1] * ip2[i2];
    int64_t src = ((t * N0 + i0) * N1 + i1) * N2 + i2;
    if (REAL_IN) {
      v.x = f[src] * s;
    } else {
      v.x = f[2 * src] * s;
      v.y = f[2 * src + 1] * s;
    }
  }
  fine[idx] = v;
}

// Fine grid -> deconvolved modes (one thread per output mode).
template <typename T, bool REAL_OUT>
__global__ void k_deconv_crop(const Cplx<T>* __restrict__ fine, T* __restrict__ f, int64_t ntrans,
                              int N0, int N1, int N2, int nf0, int nf1, int nf2,
                              const T* __restrict__ ip0, const T* __restrict__ ip1,
                              const T* __restrict__ ip2) {
  const int64_t nm = (int64_t)N0 * N1 * N2;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nm * ntrans) return;
  int64_t t = idx / nm;
  int64_t r = idx - t * nm;
  int i2 = (int)(r % N2);
  int64_t r2 = r / N2;
  int i1 = (int)(r2 % N1);
  int i0 = (int)(r2 / N1);
  auto fine_index = [](int i, int N, int nf) -> int {
    int n = i - N / 2;
    return n >= 0 ? n : n + nf;
  };
  int m0 = fine_index(i0, N0, nf0), m1 = fine_index(i1, N1, nf1), m2 = fine_index(i2, N2, nf2);
  Cplx<T> v = fine[((t * nf0 + m0) * (int64_t)nf1 + m1) * nf2 + m2];
  T s = ip0[i0] * ip1[i1] * ip2[i2];
  if (REAL_OUT) {
    f[idx] = v.x * s;
  } else {
    f[2 * idx] = v.x * s;
    f[2 * idx + 1] = v.y * s;
  }
}


// Point-order permutations between the caller's layout c[t][j] (j in input
// order) and the sorted layouts the bin kernels use.  Both walk j in input
// order so that the c[t][j] side is coalesced; the sorted side is a
// contiguous run per point.
//   to_sorted:   cs[spos[j]][t]        = c[t][j]      (spreading input, row stride ldc >= ntrans,
//                                                      zero past ntrans)
//   from_sorted: c[t][j] = cs[t / TC][spos[j]][t % TC] (interpolation output)
// Real volumes (type 2): modes -> zero-padded, deconvolved REAL fine grid, the input of a
// real-to-complex FFT that writes only the Hermitian half.  One wavefront per fine-grid row
// (t, m0, m1) writes the nf2 values of the row with 16-byte stores (nf2 is a multiple of 8);
// rows outside the modes of axes 0/1 are zero.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_deconv_pad_r(const T* __restrict__ f, T* __restrict__ fine_r, int64_t ntrans, int64_t nrun, int N0, int N1,
                   int N2, int nf0, int nf1, int nf2, const T* __restrict__ ip0, const T* __restrict__ ip1,
                   const T* __restrict__ ip2) {
  const int64_t row = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int64_t nrows = ntrans * nf0 * nf1;
  if (row >= nrows) return;
  const int m1 = (int)(row % nf1);
  const int64_t r1 = row / nf1;
  const int m0 = (int)(r1 % nf0);
  const int64_t t = r1 / nf0;
  auto mode_index = [](int m, int N, int nf) -> int {
    int hp = (N + 1) / 2, hn = N / 2;
    if (m < hp) return m + hn;
    if (m >= nf - hn) return m - nf + hn;
    return -1;
  };
  const int i0 = mode_index(m0, N0, nf0), i1 = mode_index(m1, N1, nf1);
  const bool live = t < nrun && i0 >= 0 && i1 >= 0;
  const T s01 = live ? ip0[i0] * ip1[i1] : (T)0;
  const T* src = f + ((t * N0 + (live ? i0 : 0)) * N1 + (live ? i1 : 0)) * N2;
  T* dst = fine_r + row * nf2;
  for (int m = 4 * lane; m < nf2; m += 4 * 64) {
    T v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int i2 = mode_index(m + k, N2, nf2);
      v[k] = (live && i2 >= 0) ? src[i2] * (s01 * ip2[i2]) : (T)0;
    }
    if constexpr (sizeof(T) == 4) {
      *reinterpret_cast<float4*>(dst + m) = float4{v[0], v[1], v[2], v[3]};
    } else {
      reinterpret_cast<double2*>(dst + m)[0] = double2{v[0], v[1]};
      reinterpret_cast<double2*>(dst + m)[1] = double2{v[2], v[3]};
    }
  }
}

// Real outputs (type 1): Re(F(G)) is the complex-to-real transform of the Hermitian part of the
// spread grid, H[m] = (G[m] + conj(G[-m])) / 2, of which only m2 <= nf2/2 is stored
// (conjugated when the transform sign is negative; one thread per half-grid cell).
template <typename T>
__global__ void k_fold_half(const Cplx<T>* __restrict__ G, Cplx<T>* __restrict__ H, int64_t ntrans, int nf0,
                            int nf1, int nf2, int conj_out) {
  const int h2 = nf2 / 2 + 1;
  const int64_t nh = (int64_t)nf0 * nf1 * h2;
  const int64_t nfine = (int64_t)nf0 * nf1 * nf2;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nh * ntrans) return;
  const int64_t t = idx / nh;
  const int64_t r = idx - t * nh;
  const int m2 = (int)(r % h2);
  const int64_t r2 = r / h2;
  const int m1 = (int)(r2 % nf1);
  const int m0 = (int)(r2 / nf1);
  const int n0 = m0 ? nf0 - m0 : 0, n1 = m1 ? nf1 - m1 : 0, n2 = m2 ? nf2 - m2 : 0;
  const Cplx<T> a = G[t * nfine + ((int64_t)m0 * nf1 + m1) * nf2 + m2];
  const Cplx<T> b = G[t * nfine + ((int64_t)n0 * nf1 + n1) * nf2 + n2];
  const T im = (T)0.5 * (a.y - b.y);
  H[idx] = Cplx<T>{(T)0.5 * (a.x + b.x), conj_out ? -im : im};
}

// Real fine grid (output of the complex-to-real transform) -> deconvolved real modes.
template <typename T>
__global__ void k_deconv_crop_r(const T* __restrict__ fine_r, T* __restrict__ f, int64_t ntrans, int N0, int N1,
                                int N2, int nf0, int nf1, int nf2, const T* __restrict__ ip0,
                                const T* __restrict__ ip1, const T* __restrict__ ip2) {
  const int64_t nm = (int64_t)N0 * N1 * N2;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nm * ntrans) return;
  const int64_t t = idx / nm;
  const int64_t r = idx - t * nm;
  const int i2 = (int)(r % N2);
  const int64_t r2 = r / N2;
  const int i1 = (int)(r2 % N1);
  const int i0 = (int)(r2 / N1);
  auto fine_index = [](int i, int N, int nf) -> int {
    int n = i - N / 2;
    return n >= 0 ? n : n + nf;
  };
  const int m0 = fine_index(i0, N0, nf0), m1 = fine_index(i1, N1, nf1), m2 = fine_index(i2, N2, nf2);
  f[idx] = fine_r[((t * nf0 + m0) * (int64_t)nf1 + m1) * nf2 + m2] * (ip0[i0] * ip1[i1] * ip2[i2]);
}

template <typename T>
__global__ void k_to_sorted(const Cplx<T>* __restrict__ c, Cplx<T>* __restrict__ cs,
                            const int* __restrict__ spos, int64_t M, int ntrans, int ldc) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  Cplx<T>* dst = cs + (int64_t)spos[j] * ldc;
  if constexpr (sizeof(T) == 4) {
    if ((ldc & 1) == 0) {
      // FP32 rows of even length: transform pairs as 16-byte stores (rows are 16-byte aligned)
      float4* d4 = reinterpret_cast<float4*>(dst);
      for (int t = 0; t < ldc; t += 2) {
        const Cplx<T> a = t < ntrans ? c[(int64_t)t * M + j] : Cplx<T>{0.f, 0.f};
        const Cplx<T> b = t + 1 < ntrans ? c[(int64_t)(t + 1) * M + j] : Cplx<T>{0.f, 0.f};
        d4[t >> 1] = float4{a.x, a.y, b.x, b.y};
      }
      return;
    }
  }
  for (int t = 0; t < ntrans; ++t) dst[t] = c[(int64_t)t * M + j];
  for (int t = ntrans; t < ldc; ++t) dst[t] = Cplx<T>{(T)0, (T)0};
}

// FP32 point-major staging for rows of even length ldc <= kStageLdc: a workgroup takes 256
// consecutive points, reads their values coalesced ([t][j]) into LDS, then writes each point's
// row as consecutive 16-byte pieces, so one wave instruction covers about 6 rows instead of
// 64 scattered ones.
constexpr int kStageLdc = 24;
__global__ void __launch_bounds__(kBlock)
    k_to_sorted_rows(const float2* __restrict__ c, float2* __restrict__ cs, const int* __restrict__ spos,
                     int64_t M, int ntrans, int ldc) {
  __shared__ float2 s_v[kBlock * (kStageLdc + 1)];
  __shared__ int s_pos[kBlock];
  const int64_t j0 = (int64_t)blockIdx.x * kBlock;
  const int tid = threadIdx.x;
  const int64_t j = j0 + tid;
  s_pos[tid] = j < M ? spos[j] : -1;
  for (int t = 0; t < ldc; ++t)
    s_v[tid * (kStageLdc + 1) + t] = (j < M && t < ntrans) ? c[(int64_t)t * M + j] : make_float2(0.f, 0.f);
  __syncthreads();
  const int hp = ldc >> 1;  // 16-byte pieces per row
  for (int e = tid; e < kBlock * hp; e += kBlock) {
    const int q = e / hp, h = e - q * hp;
    const int sp = s_pos[q];
    if (sp >= 0) {
      const float2 a = s_v[q * (kStageLdc + 1) + 2 * h], b = s_v[q * (kStageLdc + 1) + 2 * h + 1];
      reinterpret_cast<float4*>(cs + (int64_t)sp * ldc)[h] = float4{a.x, a.y, b.x, b.y};
    }
  }
}

struct SubDesc {
  const int* sub_sb;
  const int* sub_pt;
  const int* sub_n;
  const int* nsub;
};

// Subproblem of workgroup bid.  The 8 XCDs receive workgroups round-robin
// (bid % 8); the subproblem list is dealt to them in chunks of kSubChunk
// consecutive subproblems (chunk c to XCD c % 8), so each XCD walks runs of
// neighbouring super-bins (halo reuse in its L2) while the dense super-bins
// around the origin, and the empty tail of the grid (the grid is the
// subproblem capacity), are shared evenly.  Contiguous runs per XCD measured
// 1.7x more work on the busiest XCD than the mean.  Grids that are not a
// multiple of 8 kSubChunk keep the plain order.
constexpr int kSubChunk = 64;
__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
  if (nblk % (8 * kSubChunk) != 0) return bid;
  const int x = bid & 7, k = bid >> 3;
  return ((k / kSubChunk) * 8 + x) * kSubChunk + (k % kSubChunk);
}

// Bin of gather workgroup bid.  Workgroups go to the 8 XCDs round-robin
// (bid % 8), so XCD x runs bids x, x + 8, ... in order.  The bins are grouped
// in 8^3-bin tiles; each XCD walks whole tiles, b0-slowest inside a tile, so
// its few hundred resident workgroups cover a compact block whose
// neighbourhood fits in that XCD's L2.  Tile (t0, t1, t2) goes to XCD
// (t0 + 2 t1 + 4 t2) mod 8: the point density peaks around the origin, whose
// 8 wrapped corner tiles then land on 8 different XCDs (contiguous slabs per
// XCD left two XCDs with 40 % of the work).  Grids whose bin counts per axis
// are not multiples of 8, or whose tile count along axis 0 is not a multiple
// of 8, keep the plain order.
__device__ __forceinline__ void gather_bin(int bid, int nblk, const BinGeom& g, int& b0, int& b1, int& b2) {
  constexpr int TB = 8;
  const int T0 = g.nb[0] / TB, T1 = g.nb[1] / TB, T2 = g.nb[2] / TB;
  const int ntile = T0 * T1 * T2;
  if (g.nb[0] % TB || g.nb[1] % TB || g.nb[2] % TB || T0 % 8 || nblk != ntile * TB * TB * TB) {
    bin_coords(g, bid, b0, b1, b2);
    return;
  }
  const int x = bid & 7, k = bid >> 3;
  // tile ordinal on this XCD; with an order table the tiles nearest the grid centre (the ones
  // holding the most points: every slice passes through the centre) go first, so the light outer
  // tiles fill the end of the launch instead of the heavy ones
  const int jpos = k / (TB * TB * TB);
  const int j = g.xcd_tiles ? g.xcd_tiles[x * (ntile / 8) + jpos] : jpos;
  const int i = k % (TB * TB * TB);
  const int m = j / (T1 * T2), rr = j % (T1 * T2);
  const int t1 = rr / T2, t2 = rr % T2;
  const int t0 = ((x - 2 * t1 - 4 * t2) & 7) + 8 * m;
  b0 = t0 * TB + i / (TB * TB);
  b1 = t1 * TB + (i / TB) % TB;
  b2 = t2 * TB + i % TB;
}

// W kernel weights of a point at coordinate u; FP32 uses the native exp/sqrt.
template <typename T, int W>
__device__ __forceinline__ void stencil_weights(T u, T c, T beta, int& l0, T* w) {
  const T halfw = (T)(0.5 * W);
  int j0 = (int)ceil(u - halfw);
  if constexpr (sizeof(T) == 4) {
    const float beta_l2 = beta * 1.4426950408889634f;
#pragma unroll
    for (int a = 0; a < W; ++a) w[a] = es_eval_native((float)(j0 + a) - u, c, beta_l2);
  } else {
#pragma unroll
    for (int a = 0; a < W; ++a) w[a] = es_eval<T>((T)(j0 + a) - u, c, beta);
  }
  l0 = j0 + W / 2;
}

// Stencil weights from the polynomial form (EsPoly); l0 as stencil_weights.  y is exact: j0 - u
// is a multiple of u's ulp below 2^23 of them.
template <int W>
__device__ __forceinline__ void stencil_weights_poly(float u, const EsPoly& P, int& l0, float* w) {
  static_assert(W <= 8, "four mirrored tap pairs at most");
  const float halfw = 0.5f * (float)W;
  const int j0 = (int)ceilf(u - halfw);
  const float y = ((float)j0 - u + halfw) - 0.5f;
  const float y2 = y * y;
#pragma unroll
  for (int a = 0; a < (W + 1) / 2; ++a) {
    float e = P.e[a][4];
#pragma unroll
    for (int d = 3; d >= 0; --d) e = fmaf(e, y2, P.e[a][d]);
    if (2 * a + 1 == W) {
      w[a] = e;
    } else {
      float o = P.o[a][3];
#pragma unroll
      for (int d = 2; d >= 0; --d) o = fmaf(o, y2, P.o[a][d]);
      w[a] = fmaf(y, o, e);
      w[W - 1 - a] = fmaf(-y, o, e);
    }
  }
  l0 = j0 + W / 2;
}

// Element i of a buffer that no kernel writes while the caller runs, through the constant address
// space: with a wave-uniform index this is a scalar load, which does not wait for the wave's vector
// loads in flight.
__device__ __forceinline__ int load_invariant(const int* p, int i) {
  return ((const __attribute__((address_space(4))) int*)(uintptr_t)p)[i];
}

// Fine-grid index of an unwrapped coordinate within one period of [0, n).
__device__ __forceinline__ int wrap1(int m, int n) {
  m += m < 0 ? n : 0;
  return m >= n ? m - n : m;
}

// Stages the (8 + W)^3 halo tile of one super-bin, TC transforms from t0, into
// LDS as [cell][TC].  Loads are issued kStageGroup iterations at a time into
// registers before any is stored, so a thread has up to kStageGroup * TC
// global loads in flight instead of one iteration's.
constexpr int kStageGroup = 7;
// k_interp, whose occupancy is bounded by its tile, issues a whole W = 7 tile (14 iterations) at once
constexpr int kInterpStageGroup = 14;

// half_h2 > 0: the fine grid holds only the Hermitian half m2 < half_h2 (= nf2/2 + 1) of the
// transform of a real volume; a cell m with m2 >= half_h2 is conj(F[-m]).  conj_all conjugates
// every value (the half grid was computed with the opposite sign).
// Plane stride (cells) of the interpolation tile [e0][e1][e2].  The FP32 tile of two transforms
// per 16-byte cell with W = 7 pads its planes from 225 to 234 cells: a ds_read_b128 from 16
// lanes is conflict-free when their cells differ mod 16, and on the slices of a central-slice
// batch the padded stride spreads the points' base cells over the 16 classes more evenly (a
// conflict model over L = 128, 1024-image subproblems: 1.99 -> 1.49 cycles per 16-lane read;
// measured conflicts were half of the LDS-active cycles).
template <typename T, int W, int TC>
constexpr int interp_plane() {
  return (sizeof(T) == 4 && TC == 2 && W == 7) ? 234 : (kSuperEdge + W) * (kSuperEdge + W);
}

// Fine-grid offset of grid cell (m0, m1, m2) (each in [0, nf)); neg: the value is the conjugate
// of the stored one (the Hermitian half grid stores m2 < half_h2 only; conj_all flips every value).
__device__ __forceinline__ int64_t fine_cell(const BinGeom& g, int m0, int m1, int m2, int half_h2, int conj_all,
                                            bool& neg) {
  neg = false;
  if (!half_h2) return ((int64_t)m0 * g.nf[1] + m1) * g.nf[2] + m2;
  if (m2 >= half_h2) {
    m0 = m0 ? g.nf[0] - m0 : 0;
    m1 = m1 ? g.nf[1] - m1 : 0;
    m2 = g.nf[2] - m2;
    neg = true;
  }
  neg = neg != (conj_all != 0);
  return ((int64_t)m0 * g.nf[1] + m1) * half_h2 + m2;
}

template <typename T, int W, int TC, int PL = (kSuperEdge + W) * (kSuperEdge + W), int SG = kStageGroup>
__device__ __forceinline__ void stage_tile(const Cplx<T>* __restrict__ fine, Cplx<T>* tile, const BinGeom& g,
                                           int o0, int o1, int o2, int t0, int ntrans, int half_h2 = 0,
                                           int conj_all = 0) {
  constexpr int E = kSuperEdge + W;
  static_assert(PL >= E * E, "tile plane stride below the plane size");
  constexpr int ncell = E * E * E;
  constexpr int NIT = (ncell + kBlock - 1) / kBlock;
  const int64_t nfine = half_h2 ? (int64_t)g.nf[0] * g.nf[1] * half_h2 : (int64_t)g.nf[0] * g.nf[1] * g.nf[2];
#pragma unroll
  for (int it0 = 0; it0 < NIT; it0 += SG) {
    Cplx<T> v[SG][TC];
#pragma unroll
    for (int k = 0; k < SG; ++k) {
      const int cidx = threadIdx.x + (it0 + k) * kBlock;
      if (it0 + k < NIT && cidx < ncell) {
        const int e2 = cidx % E;
        const int e1 = (cidx / E) % E;
        const int e0 = cidx / (E * E);
        bool neg;
        const int64_t off = fine_cell(g, wrap1(o0 + e0, g.nf[0]), wrap1(o1 + e1, g.nf[1]), wrap1(o2 + e2, g.nf[2]),
                                      half_h2, conj_all, neg);
#pragma unroll
        for (int tt = 0; tt < TC; ++tt) {
          Cplx<T> x = t0 + tt < ntrans ? fine[(t0 + tt) * nfine + off] : Cplx<T>{(T)0, (T)0};
          if (neg) x.y = -x.y;
          v[k][tt] = x;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < SG; ++k) {
      const int cidx = threadIdx.x + (it0 + k) * kBlock;
      if (it0 + k < NIT && cidx < ncell) {
        const int tidx = cidx + (cidx / (E * E)) * (PL - E * E);
#pragma unroll
        for (int tt = 0; tt < TC; ++tt) tile[tidx * TC + tt] = v[k][tt];
      }
    }
  }
}

// Bank-aware visiting order of the points of each interpolation subproblem (FP32 tiles of two
// transforms per 16-byte cell, W <= 8).  ds_read_b128 serves 16 lanes per clock from 64 banks
// (measured), so a read is conflict-free when the 16 lanes' cells differ mod 16.  The interpolation
// kernels hold one point per lane pair (interp_pair_f32): a 16-lane read covers 8 consecutive
// slots, whose two halves read cells 4 planes apart (8 mod 16 at W = 7), so it is conflict-free
// when the 8 points' base cells differ mod 8.  Points are bucketed by base cell mod 8 and dealt
// round-robin over the buckets: bucket b, rank r goes to slot sum_b' min(c_b', r) +
// #{b' < b : c_b' > r}.  The point at slot s is stored as a record irec[p0 + s] = (its three
// grid coordinates, its sorted index as int bits) with its caller index iidx[p0 + s] = sidx[p]:
// k_interp then reads its points with coalesced loads that depend on nothing, issued before the
// tile is staged.  The order depends on the points only, so it is computed once per setpts.
template <int W>
__global__ void __launch_bounds__(kBlock)
    k_bank_order(SubDesc sd, const float* __restrict__ sv0, const float* __restrict__ sv1,
                 const float* __restrict__ sv2, const int* __restrict__ sidx, float4* __restrict__ irec,
                 int* __restrict__ iidx) {
  const int sub = blockIdx.x;
  if (sub >= *sd.nsub) return;
  constexpr int E = kSuperEdge + W;
  constexpr int kPer = kMaxSub / kBlock;
  constexpr int NB = 8;  // bank classes
  __shared__ int s_cnt[NB];
  const int p0 = sd.sub_pt[sub];
  const int np = sd.sub_n[sub];
  if (threadIdx.x < NB) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const float halfw = 0.5f * (float)W;
  int rb[kPer], rr[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int slot = threadIdx.x + k * kBlock;
    if (slot < np) {
      const int p = p0 + slot;
      const int m0 = min(max((int)ceilf(sv0[p] - halfw) + W / 2, 0), kSuperEdge);
      const int m1 = min(max((int)ceilf(sv1[p] - halfw) + W / 2, 0), kSuperEdge);
      const int m2 = min(max((int)ceilf(sv2[p] - halfw) + W / 2, 0), kSuperEdge);
      rb[k] = (m0 * interp_plane<float, W, 2>() + m1 * E + m2) % NB;
      rr[k] = atomicAdd(&s_cnt[rb[k]], 1);
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int slot = threadIdx.x + k * kBlock;
    if (slot < np) {
      int pos = 0;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int cb = s_cnt[b];
        pos += min(cb, rr[k]) + ((b < rb[k] && cb > rr[k]) ? 1 : 0);
      }
      const int ps = p0 + slot;
      irec[p0 + pos] = make_float4(sv0[ps], sv1[ps], sv2[ps], __int_as_float(ps));
      iidx[p0 + pos] = sidx[ps];
    }
  }
}

// Taps of one point on a staged FP32 tile of TC/2 16-byte words per cell (TC even, W <= 8):
// ax/ay[tt] += sum over the W^3 stencil of w0 w1 w2 * tile value of transform tt.  Stencil rows
// (a0, a1) go in order, double-buffered: the W cells of the next row are read from LDS while the
// FMAs of the current row run (occupancy is bounded by the tile, so registers are spent on reads
// in flight).  Each row's W taps are weighted along z first, then by the row weight w0 w1 once.
// The FP32 interpolation kernels with point records use interp_pair_f32 instead.
template <int W, int TC, int PL>
__device__ __forceinline__ void interp_taps_f32(const float4* __restrict__ t4, int l0, int l1, int l2,
                                                const float* w0, const float* w1, const float* w2, float* ax,
                                                float* ay) {
  constexpr int E = kSuperEdge + W;
  constexpr int H = TC / 2;
  float4 buf[2][W * H];
  auto load_row = [&](int a0, int a1, float4* b) {
    const int row = (l0 + a0) * PL + (l1 + a1) * E + l2;
#pragma unroll
    for (int a2 = 0; a2 < W; ++a2)
#pragma unroll
      for (int h = 0; h < H; ++h) b[a2 * H + h] = t4[(row + a2) * H + h];
  };
  load_row(0, 0, buf[0]);
#pragma unroll
  for (int k = 0; k < W * W; ++k) {
    const int a0 = k / W, a1 = k % W;
    if (k + 1 < W * W) load_row((k + 1) / W, (k + 1) % W, buf[(k + 1) & 1]);
    float rx[TC], ry[TC];
#pragma unroll
    for (int h = 0; h < H; ++h) {
      const float4 v = buf[k & 1][h];
      rx[2 * h] = w2[0] * v.x;
      ry[2 * h] = w2[0] * v.y;
      rx[2 * h + 1] = w2[0] * v.z;
      ry[2 * h + 1] = w2[0] * v.w;
    }
#pragma unroll
    for (int a2 = 1; a2 < W; ++a2) {
#pragma unroll
      for (int h = 0; h < H; ++h) {
        const float4 v = buf[k & 1][a2 * H + h];
        rx[2 * h] += w2[a2] * v.x;
        ry[2 * h] += w2[a2] * v.y;
        rx[2 * h + 1] += w2[a2] * v.z;
        ry[2 * h + 1] += w2[a2] * v.w;
      }
    }
    const float w01 = w0[a0] * w1[a1];
#pragma unroll
    for (int tt = 0; tt < TC; ++tt) {
      ax[tt] += w01 * rx[tt];
      ay[tt] += w01 * ry[tt];
    }
  }
}

// Value of lane (lane ^ 1) of the wave (a quad_perm [1, 0, 3, 2] DPP move).
__device__ __forceinline__ float swap_pair(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0xB1, 0xF, 0xF, false));
}

// Taps of one point on the FP32 tile of two transforms per 16-byte cell (W <= 8), split over a lane
// pair: lanes 2k and 2k + 1 of a wave hold the same point, and lane half h = lane & 1 runs half of
// its W^2 stencil rows.  Half 0 takes axis-0 planes [0, W/2), half 1 planes [W - W/2, W); for odd W
// the middle plane W/2 is shared, rows [0, C) to half 0 and [C, W) to half 1 (C = (W + 1) / 2).  Half
// 1 walks the middle plane from row C - 1, which belongs to half 0, with weight 0, so both halves run
// C rows at the same static offsets and every read stays inside the point's stencil.  Each plane's
// rows go in order with the next row's reads in flight, are summed along z and then weighted by
// w0 w1, as in interp_taps_f32 (the planes of a half are a loop, which keeps the compiler from
// hoisting later rows' reads).  The two halves' sums are added across the pair (one DPP swap per
// value) and half h returns transform h.  Against one thread per point this halves the lanes a
// short subproblem leaves idle, and a 16-lane LDS read covers 8 points whose halves sit 4 planes
// (8 cells mod 16 at W = 7) apart, so the bank-aware visiting order needs the 8 points' base cells
// distinct mod 8 only.
template <int W, int PL>
__device__ __forceinline__ float2 interp_pair_f32(const float4* __restrict__ t4, int h, int l0, int l1, int l2,
                                                  const float* wp, float wmid, const float* w1, const float* w2) {
  static_assert(W <= 8, "one 16-byte word of two transforms per cell");
  constexpr int E = kSuperEdge + W;
  constexpr int FULL = W / 2;                    // whole axis-0 planes per half
  constexpr int C = (W & 1) ? (W + 1) / 2 : 0;  // middle-plane rows per half
  float ax0 = 0.f, ay0 = 0.f, ax1 = 0.f, ay1 = 0.f;
  // n rows of one plane from LDS cell pa: row a1 at pa + a1 E, weight wp * wr[a1]
  auto plane = [&](int pa, float wp, const float* wr, auto nrows) {
    constexpr int n = decltype(nrows)::value;
    float4 buf[2][W];
#pragma unroll
    for (int a2 = 0; a2 < W; ++a2) buf[0][a2] = t4[pa + a2];
#pragma unroll
    for (int a1 = 0; a1 < n; ++a1) {
      if (a1 + 1 < n) {
#pragma unroll
        for (int a2 = 0; a2 < W; ++a2) buf[(a1 + 1) & 1][a2] = t4[pa + (a1 + 1) * E + a2];
      }
      float rx0 = w2[0] * buf[a1 & 1][0].x, ry0 = w2[0] * buf[a1 & 1][0].y;
      float rx1 = w2[0] * buf[a1 & 1][0].z, ry1 = w2[0] * buf[a1 & 1][0].w;
#pragma unroll
      for (int a2 = 1; a2 < W; ++a2) {
        const float4 v = buf[a1 & 1][a2];
        rx0 += w2[a2] * v.x;
        ry0 += w2[a2] * v.y;
        rx1 += w2[a2] * v.z;
        ry1 += w2[a2] * v.w;
      }
      const float ww = wp * wr[a1];
      ax0 += ww * rx0;
      ay0 += ww * ry0;
      ax1 += ww * rx1;
      ay1 += ww * ry1;
    }
  };
  // this half's plane weights (wp[a]: plane (h ? W - FULL : 0) + a), consumed in order
  float wq[FULL > 0 ? FULL : 1];
#pragma unroll
  for (int a = 0; a < FULL; ++a) wq[a] = wp[a];
  const int base = (l0 + (h ? W - FULL : 0)) * PL + l1 * E + l2;
#pragma unroll 1
  for (int a = 0; a < FULL; ++a) {
    plane(base + a * PL, wq[0], w1, std::integral_constant<int, W>{});
#pragma unroll
    for (int q = 0; q + 1 < FULL; ++q) wq[q] = wq[q + 1];
  }
  if constexpr (C > 0) {
    float wm[C];
#pragma unroll
    for (int j = 0; j < C; ++j) wm[j] = h ? (j == 0 ? 0.f : w1[C - 1 + j]) : w1[j];
    plane((l0 + W / 2) * PL + (l1 + (h ? C - 1 : 0)) * E + l2, wmid, wm, std::integral_constant<int, C>{});
  }
  // the partner keeps the other transform: send it this half's sum of that one
  const float px = swap_pair(h ? ax0 : ax1), py = swap_pair(h ? ay0 : ay1);
  return make_float2((h ? ax1 : ax0) + px, (h ? ay1 : ay0) + py);
}

// One point of a lane pair (interp_pair_f32) from its grid coordinates: the stencil weights, then
// the taps.  With the exp/sqrt weights, half 0 evaluates the axis-1 weights and half 1 the axis-2
// weights, and the pair swaps them (one DPP move each); each half evaluates only the axis-0
// weights of its own planes.  The values are those of stencil_weights for the same taps.
template <int W, int PL>
__device__ __forceinline__ float2 interp_pair_point_f32(const float4* __restrict__ t4, int h, float u0, float u1,
                                                        float u2, float kc, float beta, const EsPoly& poly) {
  constexpr int FULL = W / 2;
  float w1[W], w2[W], wp[FULL > 0 ? FULL : 1], wmid = 0.f;
  int l0, l1, l2;
  if (poly.on) {
    float w0[W];
    stencil_weights_poly<W>(u0, poly, l0, w0);
    stencil_weights_poly<W>(u1, poly, l1, w1);
    stencil_weights_poly<W>(u2, poly, l2, w2);
#pragma unroll
    for (int a = 0; a < FULL; ++a) wp[a] = h ? w0[W - FULL + a] : w0[a];
    if constexpr ((W & 1) != 0) wmid = w0[W / 2];
  } else {
    const float halfw = 0.5f * (float)W;
    const float beta_l2 = beta * 1.4426950408889634f;
    const int j0 = (int)ceilf(u0 - halfw), j1 = (int)ceilf(u1 - halfw), j2 = (int)ceilf(u2 - halfw);
    l0 = j0 + W / 2;
    l1 = j1 + W / 2;
    l2 = j2 + W / 2;
    const float uo = h ? u2 : u1;
    const int jo = h ? j2 : j1;
#pragma unroll
    for (int a = 0; a < W; ++a) {
      const float own = es_eval_native((float)(jo + a) - uo, kc, beta_l2);
      const float oth = swap_pair(own);
      w1[a] = h ? oth : own;
      w2[a] = h ? own : oth;
    }
    const int f0 = j0 + (h ? W - FULL : 0);
#pragma unroll
    for (int a = 0; a < FULL; ++a) wp[a] = es_eval_native((float)(f0 + a) - u0, kc, beta_l2);
    if constexpr ((W & 1) != 0) wmid = es_eval_native((float)(j0 + W / 2) - u0, kc, beta_l2);
  }
  l0 = min(max(l0, 0), kSuperEdge);
  l1 = min(max(l1, 0), kSuperEdge);
  l2 = min(max(l2, 0), kSuperEdge);
  return interp_pair_f32<W, PL>(t4, h, l0, l1, l2, wp, wmid, w1, w2);
}

// Type-2 interpolation: c[t, j] = sum_taps phi * fine[t, tap].
// One workgroup per (super-bin subproblem, transform chunk).  The 8^3 tile
// plus a W-cell halo is staged in LDS as [cell][TC/2] 16-byte words holding
// two complex transforms each, so one tap of one point is TC/2 ds_read_b128.
// TC = 1 (FP64 plans) stores one complex per cell.
template <typename T, int W, int TC>
__global__ void __launch_bounds__(kBlock)
    k_interp(const Cplx<T>* __restrict__ fine, Cplx<T>* __restrict__ c, SubDesc sd, BinGeom g,
             const T* __restrict__ sv0, const T* __restrict__ sv1, const T* __restrict__ sv2,
             const int* __restrict__ sidx, int64_t M, int ntrans, T kc, T beta, int half_h2, int conj_all,
             const float4* __restrict__ irec, const int* __restrict__ iidx, Cplx<T>* __restrict__ direct,
             EsPoly poly) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cplx<T>* tile = reinterpret_cast<Cplx<T>*>(smem_raw);
  const int sub = xcd_remap(blockIdx.x, gridDim.x);
  if (sub >= *sd.nsub) return;
  const int t0 = blockIdx.y * TC;
  const int sb = sd.sub_sb[sub];
  int s0, s1, s2;
  super_coords(g, sb, s0, s1, s2);
  constexpr int E = kSuperEdge + W;
  constexpr int ncell = E * E * E;
  const int o0 = s0 * kSuperEdge - W / 2, o1 = s1 * kSuperEdge - W / 2, o2 = s2 * kSuperEdge - W / 2;
  const int p0 = sd.sub_pt[sub];
  const int np = sd.sub_n[sub];
  // halo cells lie within one period of the grid (nf >= 2W), so one conditional wrap suffices
  constexpr int PL = interp_plane<T, W, TC>();
  stage_tile<T, W, TC, PL, kInterpStageGroup>(fine, tile, g, o0, o1, o2, t0, ntrans, half_h2, conj_all);
  __syncthreads();
  if constexpr (sizeof(T) == 4 && TC == 2 && W <= 8) {
    if (irec) {
      // point records in the bank-aware order of k_bank_order, one point per lane pair
      const int h = threadIdx.x & 1;
      for (int slot = threadIdx.x >> 1; slot < np; slot += kBlock / 2) {
        const float4 rec = irec[p0 + slot];
        const int p = __float_as_int(rec.w);
        const float2 v = interp_pair_point_f32<W, PL>(reinterpret_cast<const float4*>(tile), h, rec.x, rec.y, rec.z,
                                                      kc, beta, poly);
        Cplx<T>* dst = direct ? direct + (int64_t)(t0 + h) * M + iidx[p0 + slot]
                              : c + ((int64_t)blockIdx.y * M + p) * TC + h;
        if (!direct || t0 + h < ntrans) store_cplx<T>(dst, v.x, v.y);
      }
      return;
    }
  }
  for (int slot = threadIdx.x; slot < np; slot += kBlock) {
    const int p = p0 + slot;
    const int64_t j = direct ? sidx[p] : 0;
    const T u0 = sv0[p], u1 = sv1[p], u2 = sv2[p];
    T w0[W], w1[W], w2[W];
    int l0, l1, l2;
    bool use_poly = false;
    if constexpr (sizeof(T) == 4 && W <= 8) use_poly = poly.on != 0;
    if (use_poly) {
      if constexpr (sizeof(T) == 4 && W <= 8) {
        stencil_weights_poly<W>(u0, poly, l0, w0);
        stencil_weights_poly<W>(u1, poly, l1, w1);
        stencil_weights_poly<W>(u2, poly, l2, w2);
      }
    } else {
      stencil_weights<T, W>(u0, kc, beta, l0, w0);
      stencil_weights<T, W>(u1, kc, beta, l1, w1);
      stencil_weights<T, W>(u2, kc, beta, l2, w2);
    }
    l0 = min(max(l0, 0), kSuperEdge);
    l1 = min(max(l1, 0), kSuperEdge);
    l2 = min(max(l2, 0), kSuperEdge);
    T ax[TC], ay[TC];
#pragma unroll
    for (int tt = 0; tt < TC; ++tt) ax[tt] = ay[tt] = (T)0;
    if constexpr (TC % 2 == 0 && sizeof(T) == 4 && W <= 8) {
      interp_taps_f32<W, TC, PL>(reinterpret_cast<const float4*>(tile), l0, l1, l2, w0, w1, w2, ax, ay);
    } else {
      for (int a0 = 0; a0 < W; ++a0) {
#pragma unroll
        for (int a1 = 0; a1 < W; ++a1) {
          const T w01 = w0[a0] * w1[a1];
          const int row = (l0 + a0) * PL + (l1 + a1) * E + l2;
#pragma unroll
          for (int a2 = 0; a2 < W; ++a2) {
            const T ww = w01 * w2[a2];
            if constexpr (TC % 2 == 0 && sizeof(T) == 4) {
              const float4* cell = reinterpret_cast<const float4*>(tile + (row + a2) * TC);
#pragma unroll
              for (int h = 0; h < TC / 2; ++h) {
                const float4 v = cell[h];
                ax[2 * h] += ww * v.x;
                ay[2 * h] += ww * v.y;
                ax[2 * h + 1] += ww * v.z;
                ay[2 * h + 1] += ww * v.w;
              }
            } else {
#pragma unroll
              for (int tt = 0; tt < TC; ++tt) {
                const Cplx<T> v = tile[(row + a2) * TC + tt];
                ax[tt] += ww * v.x;
                ay[tt] += ww * v.y;
              }
            }
          }
        }
      }
    }
    if (direct) {
      // the caller's (ntrans, M) layout at the point's original index j
#pragma unroll
      for (int tt = 0; tt < TC; ++tt)
        if (t0 + tt < ntrans) store_cplx<T>(direct + (int64_t)(t0 + tt) * M + j, ax[tt], ay[tt]);
    } else {
      // sorted, chunk-major output: cs[chunk][p][TC]
      Cplx<T>* dst = c + ((int64_t)blockIdx.y * M + p) * TC;
#pragma unroll
      for (int tt = 0; tt < TC; ++tt) dst[tt] = Cplx<T>{ax[tt], ay[tt]};
    }
  }
}

// Type-2 interpolation, persistent and software-pipelined (FP32, two transforms per 16-byte cell,
// W <= 8, point records from k_bank_order).  Same arithmetic and visiting order as k_interp; the
// grid holds as many workgroups as are resident at once (the tile bounds occupancy to two per CU),
// a multiple of 8.  The items (transform pair, subproblem), item = pair * nsub + sub, are dealt
// to the XCDs as xcd_remap deals subproblems: runs of kSubChunk consecutive items, run q to XCD
// q % 8.  Workgroup b runs on XCD b % 8 and takes that XCD's items in order from the XCD's
// counter ctr[b % 8] (zero at launch), one item ahead: while the taps of one item run, the tile
// and point records of the next are loaded into registers, so an item's staging latency hides
// behind the previous item's arithmetic, and the counter balances the uneven subproblems.
template <int W>
__global__ void __launch_bounds__(kBlock)
    k_interp_pipe(const Cplx<float>* __restrict__ fine, Cplx<float>* __restrict__ c, SubDesc sd, BinGeom g,
                  const float4* __restrict__ irec, const int* __restrict__ iidx, int64_t M, int ntrans, float kc,
                  float beta, int half_h2, int conj_all, Cplx<float>* __restrict__ direct, int* __restrict__ ctr,
                  EsPoly poly) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float4* t4 = reinterpret_cast<float4*>(smem_raw);
  constexpr int E = kSuperEdge + W;
  constexpr int PL = interp_plane<float, W, 2>();
  constexpr int kPer = kMaxSub / kBlock;
  const int nsub = *sd.nsub;
  const int64_t nitems = (int64_t)nsub * ((ntrans + 1) / 2);
  const int xcd = blockIdx.x & 7;
  const int64_t nfine = half_h2 ? (int64_t)g.nf[0] * g.nf[1] * half_h2 : (int64_t)g.nf[0] * g.nf[1] * g.nf[2];
  // the XCD's v-th item; increasing in v, so the first one past the end ends the walk
  auto item_of = [&](int64_t v) { return ((v / kSubChunk) * 8 + xcd) * kSubChunk + v % kSubChunk; };
  __shared__ int s_v[2];
  __shared__ float4 s_rec[kPer][kBlock];
  __shared__ int s_idx[kPer][kBlock];
  if (threadIdx.x == 0) {
    s_v[0] = atomicAdd(&ctr[xcd], 1);
    s_v[1] = atomicAdd(&ctr[xcd], 1);
  }
  __syncthreads();
  // the item indices come from LDS; readfirstlane makes them wave-uniform, so the subproblem
  // descriptors are read with scalar loads (a vector load there would wait for every load in flight)
  int64_t item = item_of(__builtin_amdgcn_readfirstlane(s_v[0]));
  if (item >= nitems) return;
  // Tile staging: thread c < E^2 owns the tile column (e1, e2) = (c / E, c % E) and loads its E
  // cells e0 = 0 .. E-1; the fine-grid row of e0 is wave-uniform, so a cell costs a select and
  // an add beyond its loads.  The threads past E^2 load the last column too and store nothing:
  // with no branch around the loads, nothing joins their registers before the LDS write.
  static_assert(E * E <= kBlock, "one tile column per thread");
  const int col = threadIdx.x;
  const bool has_col = col < E * E;
  const int ce1 = min(col, E * E - 1) / E, ce2 = min(col, E * E - 1) % E;
  // next tile: cells (e0, ce1, ce2) of transforms t0 (tv0) and t0 + 1 (tv1), as loaded; the two
  // transforms stay apart until the LDS write so that no register move consumes a load early
  float2 tv0[E], tv1[E];
  bool tneg = false, ttwo = true;  // ... to be conjugated / with no second transform (zeroed)
  float4 nrec[kPer];
  int nidx[kPer];
  // records first: waiting for the tile's loads then covers them too
  auto fetch = [&](int64_t it) {
    const int sub = __builtin_amdgcn_readfirstlane((int)(it % nsub));
    const int t0 = __builtin_amdgcn_readfirstlane((int)(it / nsub)) * 2;
    const int p0 = load_invariant(sd.sub_pt, sub);
    const int np = load_invariant(sd.sub_n, sub);
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      // slots past the subproblem (never read back) repeat its last record: no branch around the loads
      // (copied by component: a whole-float4 copy here left nrec in scratch memory)
      const int slot = min((int)threadIdx.x + k * kBlock, np - 1);
      const float4 r = irec[p0 + slot];
      nrec[k].x = r.x;
      nrec[k].y = r.y;
      nrec[k].z = r.z;
      nrec[k].w = r.w;
      nidx[k] = iidx[p0 + slot];
    }
    const int sb = load_invariant(sd.sub_sb, sub);
    int s0, s1, s2;
    super_coords(g, sb, s0, s1, s2);
    const int o0 = s0 * kSuperEdge - W / 2, o1 = s1 * kSuperEdge - W / 2, o2 = s2 * kSuperEdge - W / 2;
    const bool two = t0 + 1 < ntrans;
    // the column's offset within a fine-grid x row, and whether its values are mirrored
    // (fine_cell's Hermitian-half rule: m2 >= half_h2 -> all three indices negated, conjugate)
    const int m1 = wrap1(o1 + ce1, g.nf[1]), m2 = wrap1(o2 + ce2, g.nf[2]);
    const bool mir = half_h2 && m2 >= half_h2;
    const bool neg = half_h2 && (mir != (conj_all != 0));
    const int ld2 = half_h2 ? half_h2 : g.nf[2];
    const int colo = mir ? (m1 ? g.nf[1] - m1 : 0) * ld2 + (g.nf[2] - m2) : m1 * ld2 + m2;
    const int rowst = g.nf[1] * ld2;
    const Cplx<float>* f0 = fine + (int64_t)t0 * nfine;
    const Cplx<float>* f1 = two ? f0 + nfine : f0;
    // the values stay as loaded until they are written to LDS (after the current item's taps):
    // a sign flip here would wait for the loads before the taps start
    tneg = neg;
    ttwo = two;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int r0 = wrap1(o0 + k, g.nf[0]);
      const int r0m = r0 ? g.nf[0] - r0 : 0;
      const uint32_t off = (uint32_t)((mir ? r0m : r0) * rowst + colo);
      tv0[k] = reinterpret_cast<const float2*>(f0)[off];
      tv1[k] = reinterpret_cast<const float2*>(f1)[off];
    }
  };
  fetch(item);
  while (true) {
    const int sub = (int)(item % nsub);
    const int pair = (int)(item / nsub);
    const int t0 = pair * 2;
    const int np = load_invariant(sd.sub_n, sub);
    __syncthreads();  // the previous item's taps are done with the tile
    // s_v[1] (the next item) was written before the barrier above and is rewritten only after
    // the barrier below
    const int64_t nxt = item_of(__builtin_amdgcn_readfirstlane(s_v[1]));
    // the item after next: its index is needed only at the next iteration's fetch.  The atomic is
    // issued before the LDS writes below (its latency passes while they wait for the tile loads)
    // and before the next item's loads (a wait for it leaves those in flight), as an increment
    // that wraps at 2^32 - 1 (a plain atomicAdd is turned into a wave-aggregated form that waits
    // for its result at once).
    int v2;  // assigned and read by thread 0 only, when nxt < nitems (no join that would wait for it)
    if (threadIdx.x == 0 && nxt < nitems) v2 = (int)atomicInc((unsigned*)&ctr[xcd], 0xffffffffu);
#pragma unroll
    for (int k = 0; k < E; ++k) {
      if (has_col) {
        const float2 a = tv0[k], b = tv1[k];
        const float sy = tneg ? -a.y : a.y, sw = tneg ? -b.y : b.y;
        t4[k * PL + col] = make_float4(a.x, sy, ttwo ? b.x : 0.f, ttwo ? sw : 0.f);
      }
    }
    // the current item's records move to LDS as one flat array of slots (read after the barrier below)
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      s_rec[k][threadIdx.x] = nrec[k];
      s_idx[k][threadIdx.x] = nidx[k];
    }
    __syncthreads();
    // unconditional (the last item fetches itself again, unused): a branch here would join the
    // prefetch registers and wait for the loads before the taps
    fetch(nxt < nitems ? nxt : item);
    // one point per lane pair (interp_pair_f32); the records are read back as one flat array
    const int h = threadIdx.x & 1;
    for (int slot = threadIdx.x >> 1; slot < np; slot += kBlock / 2) {
      const float4 rec = s_rec[slot / kBlock][slot % kBlock];
      const float u0 = rec.x, u1 = rec.y, u2 = rec.z;
      const int p = __float_as_int(rec.w);
      const int64_t j = s_idx[slot / kBlock][slot % kBlock];
      const float2 v = interp_pair_point_f32<W, PL>(t4, h, u0, u1, u2, kc, beta, poly);
      // one destination and one store (stores in both arms of a branch were merged into a
      // shared 4-byte half)
      Cplx<float>* dst = direct ? direct + (int64_t)(t0 + h) * M + j : c + ((int64_t)pair * M + p) * 2 + h;
      if (!direct || t0 + h < ntrans) store_cplx<float>(dst, v.x, v.y);
    }
    if (nxt >= nitems) break;
    if (threadIdx.x == 0) s_v[1] = v2;
    item = nxt;
  }
}

// Type-1 spreading as an output-driven gather: one 64-lane workgroup owns one
// 4^3 bin of fine-grid cells (one cell per lane, TC transforms in registers)
// and walks the points of the neighbouring bins in batches of up to 64.  A
// batch is prefetched into registers (one point per lane) while the previous
// batch is consumed, then staged in LDS with its 3W kernel weights (each lane
// evaluates the weights of its own point).  Points whose W^3 support misses
// the block are skipped wave-uniformly; inside, every lane adds its cell's
// contribution.  Each cell is written exactly once: no atomics, deterministic
// summation order.
constexpr int kGatherBlock = 64;
constexpr int kGatherBatch = 32;

template <typename T, int TC>
__global__ void __launch_bounds__(kGatherBlock)
    k_spread_gather(const Cplx<T>* __restrict__ cs, Cplx<T>* __restrict__ fine, BinGeom g,
                    const int* __restrict__ bin_start, const int* __restrict__ bin_count,
                    const T* __restrict__ sv0, const T* __restrict__ sv1, const T* __restrict__ sv2,
                    int64_t M, int ntrans, int W, T kc, T beta) {
  __shared__ T s_w[kGatherBatch][3][16];
  __shared__ int s_j0[kGatherBatch][3];
  __shared__ __align__(16) Cplx<T> s_c[kGatherBatch][TC];
  const int t0 = blockIdx.y * TC;
  const int lane = threadIdx.x;
  int b0, b1, b2;
  gather_bin(blockIdx.x, gridDim.x, g, b0, b1, b2);
  const int c0 = lane >> 4, c1 = (lane >> 2) & 3, c2 = lane & 3;
  const T halfw = (T)(0.5 * W);
  // neighbour bins within reach of a W-wide support: ceil((W/2) / 4)
  const int R = (W + 2 * kBinEdge - 1) / (2 * kBinEdge);
  const int nside = 2 * R + 1;
  const int nnb = nside * nside * nside;

  // iteration state over (neighbour bin, batch)
  int nbi = 0, pnext = 0, pend = 0, nq0 = 0, nq1 = 0, nq2 = 0;
  auto next_chunk = [&](int& cp, int& cn, int& cq0, int& cq1, int& cq2) -> bool {
    while (pnext >= pend) {
      if (nbi >= nnb) return false;
      const int d0 = nbi / (nside * nside) - R, d1 = (nbi / nside) % nside - R, d2 = nbi % nside - R;
      ++nbi;
      int n0 = b0 + d0, n1 = b1 + d1, n2 = b2 + d2, sh0 = 0, sh1 = 0, sh2 = 0;
      if (n0 < 0) { n0 += g.nb[0]; sh0 = -g.nf[0]; }
      if (n0 >= g.nb[0]) { n0 -= g.nb[0]; sh0 = g.nf[0]; }
      if (n1 < 0) { n1 += g.nb[1]; sh1 = -g.nf[1]; }
      if (n1 >= g.nb[1]) { n1 -= g.nb[1]; sh1 = g.nf[1]; }
      if (n2 < 0) { n2 += g.nb[2]; sh2 = -g.nf[2]; }
      if (n2 >= g.nb[2]) { n2 -= g.nb[2]; sh2 = g.nf[2]; }
      const int nbin = bin_index(g, n0, n1, n2);
      pnext = bin_start[nbin];
      pend = pnext + bin_count[nbin];
      // first cell of this block relative to the neighbour's super-bin origin (unwrapped frame)
      nq0 = kBinEdge * b0 - (n0 >> 1) * kSuperEdge - sh0;
      nq1 = kBinEdge * b1 - (n1 >> 1) * kSuperEdge - sh1;
      nq2 = kBinEdge * b2 - (n2 >> 1) * kSuperEdge - sh2;
    }
    cp = pnext;
    cn = min(kGatherBatch, pend - pnext);
    cq0 = nq0;
    cq1 = nq1;
    cq2 = nq2;
    pnext += cn;
    return true;
  };

  T ax[TC], ay[TC];
#pragma unroll
  for (int tt = 0; tt < TC; ++tt) ax[tt] = ay[tt] = (T)0;
  T rv0 = 0, rv1 = 0, rv2 = 0;
  Cplx<T> rc[TC];
  auto prefetch = [&](int cp, int cn) {
    if (lane < cn) {
      rv0 = sv0[cp + lane];
      rv1 = sv1[cp + lane];
      rv2 = sv2[cp + lane];
      const Cplx<T>* src = cs + (int64_t)(cp + lane) * ntrans + t0;
#pragma unroll
      for (int tt = 0; tt < TC; ++tt) rc[tt] = (t0 + tt < ntrans) ? src[tt] : Cplx<T>{(T)0, (T)0};
    }
  };
  int cp, cn, cq0, cq1, cq2;
  bool have = next_chunk(cp, cn, cq0, cq1, cq2);
  if (have) prefetch(cp, cn);
  while (have) {
    __syncthreads();  // previous batch fully consumed
    if (lane < cn) {
      const T vv[3] = {rv0, rv1, rv2};
#pragma unroll
      for (int axis = 0; axis < 3; ++axis) {
        const int j0 = (int)ceil(vv[axis] - halfw);
        s_j0[lane][axis] = j0;
        for (int k = 0; k < W; ++k) s_w[lane][axis][k] = es_eval<T>((T)(j0 + k) - vv[axis], kc, beta);
      }
#pragma unroll
      for (int tt = 0; tt < TC; ++tt) s_c[lane][tt] = rc[tt];
    }
    __syncthreads();
    int np_, nn_, nq0_, nq1_, nq2_;
    const bool have_next = next_chunk(np_, nn_, nq0_, nq1_, nq2_);
    if (have_next) prefetch(np_, nn_);
    for (int q = 0; q < cn; ++q) {
      const int j00 = s_j0[q][0], j01 = s_j0[q][1], j02 = s_j0[q][2];
      // wave-uniform skip: the point's support misses this 4^3 block
      if (cq0 > j00 + W - 1 || cq0 + 3 < j00 || cq1 > j01 + W - 1 || cq1 + 3 < j01 || cq2 > j02 + W - 1 ||
          cq2 + 3 < j02)
        continue;
      const int x0 = cq0 + c0 - j00, x1 = cq1 + c1 - j01, x2 = cq2 + c2 - j02;
      const bool in = x0 >= 0 && x0 < W && x1 >= 0 && x1 < W && x2 >= 0 && x2 < W;
      const T w = in ? s_w[q][0][min(max(x0, 0), W - 1)] * s_w[q][1][min(max(x1, 0), W - 1)] *
                           s_w[q][2][min(max(x2, 0), W - 1)]
                     : (T)0;
      if constexpr (TC % 2 == 0 && sizeof(T) == 4) {
        const float4* cv4 = reinterpret_cast<const float4*>(&s_c[q][0]);
#pragma unroll
        for (int h = 0; h < TC / 2; ++h) {
          const float4 v = cv4[h];
          ax[2 * h] += w * v.x;
          ay[2 * h] += w * v.y;
          ax[2 * h + 1] += w * v.z;
          ay[2 * h + 1] += w * v.w;
        }
      } else {
#pragma unroll
        for (int tt = 0; tt < TC; ++tt) {
          const Cplx<T> cv = s_c[q][tt];
          ax[tt] += w * cv.x;
          ay[tt] += w * cv.y;
        }
      }
    }
    have = have_next;
    cp = np_;
    cn = nn_;
    cq0 = nq0_;
    cq1 = nq1_;
    cq2 = nq2_;
  }
  const int g0 = kBinEdge * b0 + c0, g1 = kBinEdge * b1 + c1, g2 = kBinEdge * b2 + c2;
  const int64_t nfine = (int64_t)g.nf[0] * g.nf[1] * g.nf[2];
  const int64_t off = ((int64_t)g0 * g.nf[1] + g1) * g.nf[2] + g2;
#pragma unroll
  for (int tt = 0; tt < TC; ++tt) {
    const int t = t0 + tt;
    if (t < ntrans) fine[t * nfine + off] = Cplx<T>{ax[tt], ay[tt]};
  }
}

__device__ __forceinline__ float lane_bcast(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// Type-1 spreading of one 4^3 bin as an FP32 matrix product on the matrix
// cores (v_mfma_f32_16x16x4_f32, full FP32 products and sums):
//   F[cell][col] = sum_p W[cell][p] V[p][col],  W[cell][p] = wx_p[c0] wy_p[c1] wz_p[c2],
// with the bin's 64 cells as 4 row blocks of 16 (c0 = block, (c1, c2) = row / 4, row % 4),
// the 2 TC real columns of the point values as NCB column blocks of 16, and the hits
// 4 per step.  One 64-lane workgroup per bin (kernel width W <= 8, so the 27
// neighbour bins hold every point that reaches it).  Lane n < 27 loads neighbour
// n's point range; the ranges are walked as one flat index space, 64 points per
// chunk, the next chunk's coordinates loaded while the current one is processed.
// Each chunk's hits are appended to an LDS list (coordinates, the neighbour's
// block offset, sorted index) and lane f of the new entries evaluates weight
// f % 12 of hit f / 12 (4 cells on each axis).  Whenever 64 hits are pending
// they are consumed 4 per step: each lane forms its A entries from 3 LDS reads and
// loads its B entry from the point's sorted value row, two steps ahead.  Columns
// past 2 TC read whatever follows the row (the buffer carries a tail of slack) and
// only feed output columns that are never stored.  Each cell
// is written once (deterministic, no atomics); a lane pair exchanges halves so
// that every store is 16 bytes.
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kHitCap = 128;
constexpr int kCsortSlack = 64;  // complex values of slack after the sorted value rows

template <int TC, bool OFF32>
__global__ void __launch_bounds__(kGatherBlock) __attribute__((amdgpu_waves_per_eu(4)))
    k_spread_mfma(const Cplx<float>* __restrict__ cs, Cplx<float>* __restrict__ fine, BinGeom g,
                  const int* __restrict__ bin_start, const int* __restrict__ bin_count,
                  const float* __restrict__ sv0, const float* __restrict__ sv1,
                  const float* __restrict__ sv2, int ntrans, int ldc, int W, float kc, float beta_l2) {
  constexpr int NC = 2 * TC;
  // Columns go in blocks of 16 on v_mfma_f32_16x16x4_f32; a tail of at most 8 columns goes in
  // groups of 4 on v_mfma_f32_4x4x1_16b_f32 (measured 10.6 against 36 cycles per instruction),
  // whose 16 blocks are (hit grp, cell row c1) with rows c2 and columns 4 of the tail, so its A
  // operand is the 16x16x4 one unchanged.  Its accumulators hold per-hit-slot partial sums,
  // added across the 4 lane groups at the end.
  constexpr int NR = NC % 16;
  constexpr bool kTail4 = NR > 0 && NR <= 8;
  constexpr int NCB = kTail4 ? NC / 16 : (NC + 15) / 16;
  constexpr int NCG = kTail4 ? (NR + 3) / 4 : 0;
  constexpr int NB = NCB + NCG;  // B registers per hit
  constexpr int kNb = 27;
  __shared__ float4 s_rec[kHitCap];  // p0, p1, p2, packed block offsets q0 | q1 << 10 | q2 << 20 (+512 each)
  // per hit: OFF32 -> the float index of its value row's chunk in cs (uint32 bits; the
  // address is formed once per hit, lane-parallel, instead of once per step), else its point
  __shared__ int s_idx[kHitCap];
  __shared__ __align__(16) float s_w[kHitCap * 12];  // per hit: wx[0..3], wy[0..3], wz[0..3]
  const int t0 = blockIdx.y * TC;
  const int lane = threadIdx.x;
  int b0, b1, b2;
  gather_bin(blockIdx.x, gridDim.x, g, b0, b1, b2);
  const float halfw = 0.5f * (float)W;
  const int grp = lane >> 4, sub = lane & 15;
  f32x4 acc[4][NCB > 0 ? NCB : 1];
  f32x4 acc4[4][NCG > 0 ? NCG : 1];
#pragma unroll
  for (int c0 = 0; c0 < 4; ++c0) {
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) acc[c0][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int cg = 0; cg < NCG; ++cg) acc4[c0][cg] = f32x4{0.f, 0.f, 0.f, 0.f};
  }

  // ---- neighbour table: lane n < 27 holds neighbour n's first sorted point, the exclusive
  // prefix of the point counts and the packed offset of this block in the neighbour's frame
  int nb_start = 0, nb_cnt = 0, nb_pack = 0;
  if (lane < kNb) {
    int n0 = b0 + lane / 9 - 1, n1 = b1 + (lane / 3) % 3 - 1, n2 = b2 + lane % 3 - 1;
    int sh0 = 0, sh1 = 0, sh2 = 0;
    if (n0 < 0) { n0 += g.nb[0]; sh0 = -g.nf[0]; }
    if (n0 >= g.nb[0]) { n0 -= g.nb[0]; sh0 = g.nf[0]; }
    if (n1 < 0) { n1 += g.nb[1]; sh1 = -g.nf[1]; }
    if (n1 >= g.nb[1]) { n1 -= g.nb[1]; sh1 = g.nf[1]; }
    if (n2 < 0) { n2 += g.nb[2]; sh2 = -g.nf[2]; }
    if (n2 >= g.nb[2]) { n2 -= g.nb[2]; sh2 = g.nf[2]; }
    const int nbin = bin_index(g, n0, n1, n2);
    nb_start = bin_start[nbin];
    nb_cnt = bin_count[nbin];
    // the block's first cell in the neighbour's super-bin frame (unwrapped across the boundary)
    const int q0 = kBinEdge * b0 - (n0 >> 1) * kSuperEdge - sh0;
    const int q1 = kBinEdge * b1 - (n1 >> 1) * kSuperEdge - sh1;
    const int q2 = kBinEdge * b2 - (n2 >> 1) * kSuperEdge - sh2;
    nb_pack = (q0 + 512) | ((q1 + 512) << 10) | ((q2 + 512) << 20);
  }
  int incl = nb_cnt;
#pragma unroll
  for (int off = 1; off < 32; off <<= 1) {
    const int t = __shfl_up(incl, off);
    if (lane >= off) incl += t;
  }
  const int nb_pre = incl - nb_cnt;
  const int total = __builtin_amdgcn_readlane(incl, kNb - 1);

  auto load_b = [&](int slot, float* bv) {
    const float* row = OFF32 ? reinterpret_cast<const float*>(cs) + (uint32_t)s_idx[slot]
                             : reinterpret_cast<const float*>(cs + (int64_t)s_idx[slot] * ldc + t0);
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) bv[cb] = row[sub + 16 * cb];
#pragma unroll
    for (int cg = 0; cg < NCG; ++cg) bv[NCB + cg] = row[16 * NCB + 4 * cg + (sub & 3)];
  };
  // A cell block (one x level of the bin) whose x weight is zero for all 4 hits of a step adds
  // nothing: its 3 MFMAs are skipped (wave-uniform ballot per level).
  auto mma = [&](int slot, const float* bv) {
    const float4 wx = *reinterpret_cast<const float4*>(&s_w[slot * 12]);
    const float wyz = s_w[slot * 12 + 4 + (sub >> 2)] * s_w[slot * 12 + 8 + (sub & 3)];
    const float wxa[4] = {wx.x, wx.y, wx.z, wx.w};
#pragma unroll
    for (int c0 = 0; c0 < 4; ++c0) {
      if (__ballot(wxa[c0] != 0.f)) {
        const float a = wxa[c0] * wyz;
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb)
          acc[c0][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bv[cb], acc[c0][cb], 0, 0, 0);
#pragma unroll
        for (int cg = 0; cg < NCG; ++cg)
          acc4[c0][cg] = __builtin_amdgcn_mfma_f32_4x4x1f32(a, bv[NCB + cg], acc4[c0][cg], 0, 0, 0);
      }
    }
  };
  // steps over hit slots [0, nfull), nfull a multiple of 8; the B entries of step s + 2 are
  // loaded during step s (two register sets, no copies between them).  The loads are issued
  // unconditionally (the last ones repeat valid slots) so that each wait covers exactly the
  // set it needs.
  auto steps = [&](int nfull) {
    float ba[NB], bb[NB];
    load_b(grp, ba);
    load_b(4 + grp, bb);
    for (int s = 0; s < nfull; s += 8) {
      const int sn = s + 8 < nfull ? s + 8 : s;
      mma(s + grp, ba);
      load_b(sn + grp, ba);
      mma(s + 4 + grp, bb);
      load_b(sn + 4 + grp, bb);
    }
  };
  // flat index f of the neighbourhood -> sorted point index and packed offset; ncur is the
  // neighbour of the chunk's first index (wave-uniform)
  auto locate = [&](int base, int ncur, int& pi, int& pk, int& nlast) {
    const int f = base + lane;
    int n = ncur;
    for (int m = ncur + 1; m < kNb; ++m) {
      const int pm = __builtin_amdgcn_readlane(nb_pre, m);
      if (pm > base + kGatherBlock - 1) break;
      if (f >= pm) n = m;
    }
    pi = __shfl(nb_start, n) + f - __shfl(nb_pre, n);
    pk = __shfl(nb_pack, n);
    nlast = __builtin_amdgcn_readlane(n, kGatherBlock - 1);
  };

  int nh = 0;  // pending hits in the list (wave-uniform)
  int ncur = 0;
  int pi_n = 0, pk_n = 0;
  float v0n = 0.f, v1n = 0.f, v2n = 0.f;
  if (total > 0) {
    locate(0, 0, pi_n, pk_n, ncur);
    if (lane < total) {
      v0n = sv0[pi_n];
      v1n = sv1[pi_n];
      v2n = sv2[pi_n];
    }
  }
  for (int base = 0; base < total; base += kGatherBlock) {
    const int pi = pi_n, pk = pk_n;
    const float v0 = v0n, v1 = v1n, v2 = v2n;
    const bool valid = base + lane < total;
    {
      // the next chunk's coordinates, loaded unconditionally (lanes past the neighbourhood, and
      // every lane after the last chunk, read point 0 and never use it): loads under a branch
      // would be joined, and waited for, before this chunk's work
      const int nb = base + kGatherBlock;
      locate(nb < total ? nb : base, ncur, pi_n, pk_n, ncur);
      const int pl = nb + lane < total ? pi_n : 0;
      v0n = sv0[pl];
      v1n = sv1[pl];
      v2n = sv2[pl];
    }
    bool hit = false;
    if (valid) {
      const int q0 = (pk & 1023) - 512, q1 = ((pk >> 10) & 1023) - 512, q2 = ((pk >> 20) & 1023) - 512;
      const int j0 = (int)ceilf(v0 - halfw), j1 = (int)ceilf(v1 - halfw), j2 = (int)ceilf(v2 - halfw);
      hit = !(q0 > j0 + W - 1 || q0 + 3 < j0 || q1 > j1 + W - 1 || q1 + 3 < j1 || q2 > j2 + W - 1 ||
              q2 + 3 < j2);
    }
    const unsigned long long mask = __ballot(hit);
    const int nnew = __popcll(mask);
    if (nnew == 0) continue;
    if (hit) {
      const int r = nh + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
      s_rec[r] = make_float4(v0, v1, v2, __int_as_float(pk));
      s_idx[r] = OFF32 ? (int)((uint32_t)pi * (uint32_t)(2 * ldc) + (uint32_t)(2 * t0)) : pi;
    }
    __syncthreads();
    // one lane per (hit, axis): the axis's 4 weights (the bin's 4 cells), stored as one float4
    for (int f = lane; f < 3 * nnew; f += kGatherBlock) {
      const int hit = f / 3, ax = f - 3 * hit;
      const int slot = nh + hit;
      const float4 rec = s_rec[slot];
      const float p = ax == 0 ? rec.x : (ax == 1 ? rec.y : rec.z);
      const int sh = ax == 0 ? 0 : (ax == 1 ? 10 : 20);
      const int qa = ((__float_as_int(rec.w) >> sh) & 1023) - 512;
      const int j0 = (int)ceilf(p - halfw);
      float w4[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int x = qa + c - j0;
        w4[c] = (x >= 0 && x < W) ? es_eval_native((float)(qa + c) - p, kc, beta_l2) : 0.f;
      }
      *reinterpret_cast<float4*>(&s_w[slot * 12 + 4 * ax]) = make_float4(w4[0], w4[1], w4[2], w4[3]);
    }
    __syncthreads();
    nh += nnew;
    if (nh >= kGatherBlock) {
      const int nfull = nh & ~7;
      steps(nfull);
      const int rem = nh - nfull;  // < 8: moved to the front of the list
      float4 rec;
      int idx = 0;
      float w[12];
      if (lane < rem) {
        rec = s_rec[nfull + lane];
        idx = s_idx[nfull + lane];
#pragma unroll
        for (int e = 0; e < 12; ++e) w[e] = s_w[(nfull + lane) * 12 + e];
      }
      __syncthreads();
      if (lane < rem) {
        s_rec[lane] = rec;
        s_idx[lane] = idx;
#pragma unroll
        for (int e = 0; e < 12; ++e) s_w[lane * 12 + e] = w[e];
      }
      __syncthreads();
      nh = rem;
    }
  }
  if (nh > 0) {
    // pad the last steps with zero-weight entries that read a valid row
    const int nfull = (nh + 7) & ~7;
    if (lane >= nh && lane < nfull) {
      s_idx[lane] = s_idx[0];
#pragma unroll
      for (int e = 0; e < 12; ++e) s_w[lane * 12 + e] = 0.f;
    }
    __syncthreads();
    steps(nfull);
  }

  // D layout: lane l, register v -> cell (c0, c1 = l / 16, c2 = v), column l % 16 (+ 16 cb).
  // Lanes 2m (real parts) and 2m + 1 (imaginary parts) of one transform swap halves, then the
  // even lane stores cells c2 = 0, 1 and the odd lane cells 2, 3 as complex pairs.
  const int64_t nfine = (int64_t)g.nf[0] * g.nf[1] * g.nf[2];
  const bool odd = sub & 1;
  const int g1 = kBinEdge * b1 + grp;
  const int g2 = kBinEdge * b2 + (odd ? 2 : 0);
#pragma unroll
  for (int cb = 0; cb < NCB; ++cb) {
    const int tt = (sub + 16 * cb) >> 1;
    const int t = t0 + tt;
#pragma unroll
    for (int c0 = 0; c0 < 4; ++c0) {
      const f32x4 d = acc[c0][cb];
      const float r0 = __shfl_xor(odd ? d[0] : d[2], 1);
      const float r1 = __shfl_xor(odd ? d[1] : d[3], 1);
      const float4 out = odd ? make_float4(r0, d[2], r1, d[3]) : make_float4(d[0], r0, d[1], r1);
      if (tt < TC && t < ntrans) {
        const int64_t off = ((int64_t)(kBinEdge * b0 + c0) * g.nf[1] + g1) * g.nf[2] + g2;
        *reinterpret_cast<float4*>(fine + t * nfine + off) = out;
      }
    }
  }
  // 4x4 tail: lane l, register r -> cell (c1 = (l / 4) % 4, c2 = r), tail column l % 4 (+ 4 cg),
  // one partial sum per lane group; after the sums over the groups, lanes l < 16 with an even
  // column write (re, im) pairs.
#pragma unroll
  for (int cg = 0; cg < NCG; ++cg) {
#pragma unroll
    for (int c0 = 0; c0 < 4; ++c0) {
      f32x4 d = acc4[c0][cg];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        d[r] += __shfl_xor(d[r], 16);
        d[r] += __shfl_xor(d[r], 32);
      }
      const int jc = lane & 3;
      const int colt = 16 * NCB + 4 * cg + jc;  // real column within the chunk
      const int tt = colt >> 1;
      const int t = t0 + tt;
      float im[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) im[r] = __shfl_xor(d[r], 1);
      if (lane < 16 && (jc & 1) == 0 && tt < TC && t < ntrans) {
        const int c1 = (lane >> 2) & 3;
        const int64_t base = t * nfine + ((int64_t)(kBinEdge * b0 + c0) * g.nf[1] + kBinEdge * b1 + c1) * g.nf[2] +
                             kBinEdge * b2;
#pragma unroll
        for (int r = 0; r < 4; ++r) fine[base + r] = Cplx<float>{d[r], im[r]};
      }
    }
  }
}

// ----------------------------------------------------------------------------
// Plan
// ----------------------------------------------------------------------------

struct Plan {
  int type = 2;   // 0 = both, 1, 2
  int dtype = 0;  // 0 = f32, 1 = f64
  int iflag2 = -1;
  int64_t N[3] = {0, 0, 0};
  int64_t nf[3] = {0, 0, 0};
  int64_t ntrans = 1;
  int w = 7;
  int tc = 2;  // transforms per interpolation workgroup
  int interp_pipe = 1;    // FP32 with point records: k_interp_pipe instead of k_interp
  int pipe_grid = 0;      // k_interp_pipe's grid: resident workgroups on the device, a multiple of 8
  int* pipe_ctr = nullptr;  // k_interp_pipe's per-XCD item counters (8 ints)
  EsPoly poly{};            // polynomial stencil weights of the FP32 interpolation kernels
  bool presorted = false;  // the current type-1 call's values are already in csort (no permute)
  int max_tc_gather = 20;  // transforms per gather workgroup
  double beta = 0, eps = 1e-6;
  BinGeom geom;
  int64_t nbins = 0, nsuper = 0;
  void* fine = nullptr;
  void* invp[3] = {nullptr, nullptr, nullptr};
  int64_t M = 0, M_cap = 0;
  int64_t sub_cap = 0, sub_alloc = 0;
  int* bin_count = nullptr;
  int* bin_start = nullptr;
  int* sub_start = nullptr;
  int* sub_cnt = nullptr;     // subproblems per super-bin (scanned into sub_start)
  int* sub_sb = nullptr;
  int* sub_pt = nullptr;
  int* sub_n = nullptr;
  int* nsub = nullptr;
  int* pt_bin = nullptr;
  void* pv = nullptr;  // setpts: per-point (coordinates relative to the super-bin, pad), 4 x T
  int* pt_pos = nullptr;      // sorted keys (radix sort output)
  void* sort_tmp = nullptr;   // radix sort work area
  size_t sort_tmp_bytes = 0;
  int* sidx = nullptr;
  int* spos = nullptr;
  float4* irec = nullptr;     // interpolation point records in bank-aware order (FP32, two
  int* iidx = nullptr;        // transforms per cell, W <= 8) and their caller indices
  bool irec_valid = false;
  void* csort = nullptr;  // sorted-order staging of point values (M * ldc_cap complex)
  int ldc_cap = 0;        // row capacity of csort
  void* sv[3] = {nullptr, nullptr, nullptr};
  rocfft_plan fft_fwd = nullptr, fft_bwd = nullptr;
  // real volumes: R2C for type 2, C2R of the folded Hermitian half for type 1 (COV3D_REAL_FFT=0 disables)
  rocfft_plan fft_r2c = nullptr, fft_c2r = nullptr;
  int real_fft = 1;
  void* rbuf = nullptr;       // real grid (type 2) or Hermitian half grid (type 1), allocated on first use
  size_t rbuf_bytes = 0;
  // pruned FP32 real transforms (fft_pruned.hip) on cubic power-of-two grids instead of
  // pad + R2C and fold + C2R + crop (COV3D_FFT_MODE=0 keeps rocFFT); rbuf holds their passes
  bool pruned = false;
  void* tw = nullptr;         // exp(-2 pi i k / nf), k < nf (complex64)
  int stage_half = 0;         // interpolation reads a Hermitian half grid with this m2 extent (0: full grid)
  int stage_conj = 0;         // ... conjugated
  rocfft_execution_info info = nullptr;
  void* fft_work = nullptr;
  size_t work_size = 0;
  bool pts_set = false;
};

static std::once_flag g_rocfft_once;

static int free_points(Plan* p) {
  hipFree(p->pt_bin);
  hipFree(p->pt_pos);
  hipFree(p->sort_tmp);
  p->sort_tmp = nullptr;
  p->sort_tmp_bytes = 0;
  hipFree(p->sidx);
  hipFree(p->spos);
  hipFree(p->irec);
  hipFree(p->iidx);
  p->irec = nullptr;
  p->iidx = nullptr;
  p->irec_valid = false;
  hipFree(p->csort);
  for (int k = 0; k < 3; ++k) hipFree(p->sv[k]);
  hipFree(p->pv);
  p->pv = nullptr;
  hipFree(p->sub_sb);
  hipFree(p->sub_pt);
  hipFree(p->sub_n);
  p->pt_bin = p->pt_pos = p->sidx = p->spos = nullptr;
  p->csort = nullptr;
  p->sub_sb = p->sub_pt = p->sub_n = nullptr;
  for (int k = 0; k < 3; ++k) p->sv[k] = nullptr;
  p->M_cap = 0;
  p->sub_alloc = 0;
  return 0;
}

#define COV3D_FFT(call)                                                  \
  do {                                                                   \
    rocfft_status s_ = (call);                                           \
    if (s_ != rocfft_status_success) {                                   \
      ::cov3d::set_error("%s failed: rocfft status %d", #call, (int)s_); \
      return COV3D_ERR_FFT;                                              \
    }                                                                    \
  } while (0)

template <typename T, int W, int TC>
static int launch_interp(Plan* p, void* c_out, int ntrans_run, hipStream_t st) {
  // k_interp's tile may pad its planes (interp_plane)
  const size_t smem_k = (size_t)(kSuperEdge + W) * interp_plane<T, W, TC>() * TC * sizeof(Cplx<T>);
  COV3D_CHECK_ARG(smem_k <= 160 * 1024, "NUFFT tile of %zu bytes exceeds LDS", smem_k);
  static bool attr_done = false;
  if (smem_k > 65536 && !attr_done) {
    COV3D_HIP(hipFuncSetAttribute((const void*)k_interp<T, W, TC>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_done = true;
  }
  SubDesc sd{p->sub_sb, p->sub_pt, p->sub_n, p->nsub};
  dim3 grid((unsigned)p->sub_cap, (unsigned)ceil_div(ntrans_run, TC));
  const T kc = (T)(4.0 / ((double)W * W));
  ScopedTiming tm(kTInterp, st);
  // both kernels write the caller's (ntrans, M) layout directly (no permute pass)
  if constexpr (sizeof(T) == 4 && W <= 8 && TC == 2) {
    if (p->interp_pipe && p->irec_valid) {
      if (p->pipe_grid == 0) {
        COV3D_HIP(hipFuncSetAttribute((const void*)k_interp_pipe<W>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)smem_k));
        int dev = 0, ncu = 0, per_cu = 0;
        COV3D_HIP(hipGetDevice(&dev));
        COV3D_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        COV3D_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_interp_pipe<W>, kBlock,
                                                               smem_k));
        p->pipe_grid = std::max(8, (ncu * std::max(per_cu, 1)) / 8 * 8);
        COV3D_HIP(hipMalloc(&p->pipe_ctr, 8 * sizeof(int)));
      }
      // the XCD counters start at zero (a kernel, not a memset node, so that the launch
      // sequence stays capturable)
      hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, p->pipe_ctr, (int64_t)8, 0);
      hipLaunchKernelGGL((k_interp_pipe<W>), dim3((unsigned)p->pipe_grid), dim3(kBlock), smem_k, st,
                         (const Cplx<float>*)p->fine, (Cplx<float>*)p->csort, sd, p->geom,
                         (const float4*)p->irec, (const int*)p->iidx, p->M, ntrans_run, (float)kc,
                         (float)p->beta, p->stage_half, p->stage_conj, (Cplx<float>*)c_out, p->pipe_ctr, p->poly);
      COV3D_LAUNCH_CHECK();
      return 0;
    }
  }
  hipLaunchKernelGGL((k_interp<T, W, TC>), grid, dim3(kBlock), smem_k, st, (const Cplx<T>*)p->fine,
                     (Cplx<T>*)p->csort, sd, p->geom, (const T*)p->sv[0], (const T*)p->sv[1],
                     (const T*)p->sv[2], p->sidx, p->M, ntrans_run, kc, (T)p->beta, p->stage_half,
                     p->stage_conj, (TC == 2 && p->irec_valid) ? (const float4*)p->irec : nullptr,
                     (TC == 2 && p->irec_valid) ? (const int*)p->iidx : nullptr, (Cplx<T>*)c_out, p->poly);
  COV3D_LAUNCH_CHECK();
  return 0;
}

template <typename T, int TC>
static int dispatch_interp_w(Plan* p, void* c_out, int ntrans_run, hipStream_t st) {
  switch (p->w) {
#define COV3D_W(WW) \
  case WW:          \
    return launch_interp<T, WW, TC>(p, c_out, ntrans_run, st);
    COV3D_W(2) COV3D_W(3) COV3D_W(4) COV3D_W(5) COV3D_W(6) COV3D_W(7) COV3D_W(8) COV3D_W(9)
    COV3D_W(10) COV3D_W(11) COV3D_W(12) COV3D_W(13) COV3D_W(14) COV3D_W(15) COV3D_W(16)
#undef COV3D_W
    default:
      set_error("unsupported kernel width %d", p->w);
      return COV3D_ERR_ARG;
  }
}

static int run_interp(Plan* p, void* c_out, int ntrans_run, hipStream_t st) {
  if (p->M == 0) return 0;
  if (p->dtype == 1) return dispatch_interp_w<double, 1>(p, c_out, ntrans_run, st);
  return dispatch_interp_w<float, 2>(p, c_out, ntrans_run, st);
}

// FP32 plans spread on the matrix cores (kernel widths up to 8, which covers every FP32 tolerance:
// eps >= 10 * FLT_EPSILON gives w <= 7); FP64 plans use the LDS-staged gather.
static bool spread_on_mfma(const Plan* p) { return p->dtype == 0 && p->w <= 8; }

template <typename T, int TC>
static int launch_spread(Plan* p, const void* c_in, int ntrans_run, hipStream_t st) {
  const bool mfma = spread_on_mfma(p);
  const int ldc = mfma ? (int)ceil_div(ntrans_run, TC) * TC : ntrans_run;
  COV3D_CHECK_ARG(ldc <= p->ldc_cap, "sorted row stride %d exceeds %d", ldc, p->ldc_cap);
  if (!p->presorted) {
    ScopedTiming tm(kTPermute, st);
    if (sizeof(T) == 4 && (ldc & 1) == 0 && ldc <= kStageLdc)
      hipLaunchKernelGGL(k_to_sorted_rows, dim3((unsigned)ceil_div(p->M, kBlock)), dim3(kBlock), 0, st,
                         (const float2*)c_in, (float2*)p->csort, p->spos, p->M, ntrans_run, ldc);
    else
      hipLaunchKernelGGL((k_to_sorted<T>), dim3((unsigned)ceil_div(p->M, kBlock)), dim3(kBlock), 0, st,
                         (const Cplx<T>*)c_in, (Cplx<T>*)p->csort, p->spos, p->M, ntrans_run, ldc);
    COV3D_LAUNCH_CHECK();
  }
  dim3 grid((unsigned)p->nbins, (unsigned)ceil_div(ntrans_run, TC));
  const T kc = (T)(4.0 / ((double)p->w * p->w));
  ScopedTiming tm(kTSpread, st);
  if constexpr (sizeof(T) == 4) {
    if (mfma) {
      // 32-bit row offsets when every float index of the sorted input fits in 32 bits
      const bool off32 = ((int64_t)p->M * ldc + kCsortSlack) * 2 < ((int64_t)1 << 32);
      auto k = off32 ? k_spread_mfma<TC, true> : k_spread_mfma<TC, false>;
      hipLaunchKernelGGL(k, grid, dim3(kGatherBlock), 0, st, (const Cplx<float>*)p->csort,
                         (Cplx<float>*)p->fine, p->geom, p->bin_start, p->bin_count, (const float*)p->sv[0],
                         (const float*)p->sv[1], (const float*)p->sv[2], ntrans_run, ldc, p->w, (float)kc,
                         (float)(p->beta * 1.4426950408889634));
      COV3D_LAUNCH_CHECK();
      return 0;
    }
  }
  hipLaunchKernelGGL((k_spread_gather<T, TC>), grid, dim3(kGatherBlock), 0, st, (const Cplx<T>*)p->csort,
                     (Cplx<T>*)p->fine, p->geom, p->bin_start, p->bin_count, (const T*)p->sv[0],
                     (const T*)p->sv[1], (const T*)p->sv[2], p->M, ntrans_run, p->w, kc,
                     (T)p->beta);
  COV3D_LAUNCH_CHECK();
  return 0;
}

// Transforms per gather workgroup: at most max_tc (<= 20) in registers,
// split evenly over ceil(n / max_tc) chunks, rounded up to an instantiated TC.
static int gather_tc(int n, int max_tc) {
  const int chunks = (int)ceil_div(n, max_tc);
  const int need = (int)ceil_div(n, chunks);
  for (int tc : {1, 2, 4, 8, 10, 16})
    if (need <= tc) return tc;
  return 20;
}

// Row stride of the sorted spreading input for n transforms: chunks * TC.
static int gather_ldc(int n, int max_tc) {
  const int tc = gather_tc(n, max_tc);
  return (int)ceil_div(n, tc) * tc;
}

// Row stride of the sorted values the spreading of n transforms reads (launch_spread / run_spread).
static int spread_ldc(const Plan* p, int n) {
  return spread_on_mfma(p) ? gather_ldc(n, p->max_tc_gather) : n;
}

template <typename T>
static int run_spread(Plan* p, const void* c_in, int ntrans_run, hipStream_t st) {
  switch (gather_tc(ntrans_run, p->max_tc_gather)) {
    case 1: return launch_spread<T, 1>(p, c_in, ntrans_run, st);
    case 2: return launch_spread<T, 2>(p, c_in, ntrans_run, st);
    case 4: return launch_spread<T, 4>(p, c_in, ntrans_run, st);
    case 8: return launch_spread<T, 8>(p, c_in, ntrans_run, st);
    case 10: return launch_spread<T, 10>(p, c_in, ntrans_run, st);
    case 16: return launch_spread<T, 16>(p, c_in, ntrans_run, st);
    default: return launch_spread<T, 20>(p, c_in, ntrans_run, st);
  }
}

static int env_int(const char* name, int dflt) {
  const char* v = getenv(name);
  if (!v || !*v) return dflt;
  return atoi(v);
}

static int exec_fft(Plan* p, bool forward_sign, hipStream_t st) {
  COV3D_FFT(rocfft_execution_info_set_stream(p->info, st));
  ScopedTiming tm(kTFft, st);
  void* bufs[1] = {p->fine};
  COV3D_FFT(rocfft_execute(forward_sign ? p->fft_fwd : p->fft_bwd, bufs, nullptr, p->info));
  return 0;
}

static int exec_fft_real(Plan* p, rocfft_plan plan, void* in, void* out, hipStream_t st) {
  COV3D_FFT(rocfft_execution_info_set_stream(p->info, st));
  ScopedTiming tm(kTFft, st);
  void* ib[1] = {in};
  void* ob[1] = {out};
  COV3D_FFT(rocfft_execute(plan, ib, ob, p->info));
  return 0;
}

// Real-transform resources, created on the first real execute: the R2C / C2R plans (rocFFT
// compiles kernels per plan, so plans that never see a real volume do not pay for them), the
// staging buffer, and a larger shared work buffer if these plans need one.
static int ensure_real(Plan* p) {
  if (!p->fft_r2c) {
    size_t lengths[3] = {(size_t)p->nf[2], (size_t)p->nf[1], (size_t)p->nf[0]};
    rocfft_precision prec = p->dtype == 1 ? rocfft_precision_double : rocfft_precision_single;
    if (rocfft_plan_create(&p->fft_r2c, rocfft_placement_notinplace, rocfft_transform_type_real_forward, prec, 3,
                           lengths, (size_t)p->ntrans, nullptr) != rocfft_status_success ||
        rocfft_plan_create(&p->fft_c2r, rocfft_placement_notinplace, rocfft_transform_type_real_inverse, prec, 3,
                           lengths, (size_t)p->ntrans, nullptr) != rocfft_status_success) {
      if (p->fft_r2c) rocfft_plan_destroy(p->fft_r2c);
      p->fft_r2c = nullptr;
      set_error("rocfft real-transform plan creation failed");
      return COV3D_ERR_FFT;
    }
    size_t w3 = 0, w4 = 0;
    rocfft_plan_get_work_buffer_size(p->fft_r2c, &w3);
    rocfft_plan_get_work_buffer_size(p->fft_c2r, &w4);
    const size_t need = std::max(w3, w4);
    if (need > p->work_size) {
      // the old work buffer may still be in use by queued transforms of this plan
      COV3D_HIP(hipDeviceSynchronize());
      hipFree(p->fft_work);
      p->fft_work = nullptr;
      p->work_size = 0;
      if (hipMalloc(&p->fft_work, need) != hipSuccess) {
        set_error("hipMalloc of %zu-byte FFT work buffer failed", need);
        return COV3D_ERR_HIP;
      }
      p->work_size = need;
      COV3D_FFT(rocfft_execution_info_set_work_buffer(p->info, p->fft_work, p->work_size));
    }
  }
  if (!p->rbuf && hipMalloc(&p->rbuf, p->rbuf_bytes) != hipSuccess) {
    p->rbuf = nullptr;
    set_error("hipMalloc of %zu-byte real-transform buffer failed", p->rbuf_bytes);
    return COV3D_ERR_HIP;
  }
  return 0;
}

template <typename T>
static int exec_type2(Plan* p, const void* f, bool f_real, void* c, int64_t nrun, hipStream_t st) {
  const int64_t nfine = p->nf[0] * p->nf[1] * p->nf[2];
  const int64_t total = nfine * p->ntrans;
  dim3 g1((unsigned)ceil_div(total, kBlock));
  if (f_real && p->real_fft) {
    // real volume: real fine grid -> R2C (Hermitian half) -> interpolation mirrors m2 > nf2/2
    int rc = ensure_real(p);
    if (rc) return rc;
    if constexpr (sizeof(T) == 4) {
      if (p->pruned) {
        {
          ScopedTiming tm(kTFft, st);
          rc = pruned_type2((const float*)f, p->fine, p->rbuf, p->tw, (const float*)p->invp[0],
                            (const float*)p->invp[1], (const float*)p->invp[2], (int)p->ntrans, (int)nrun,
                            (int)p->nf[0], (int)p->N[0], st);
          if (rc) return rc;
        }
        p->stage_half = (int)(p->nf[2] / 2 + 1);
        p->stage_conj = p->iflag2 > 0 ? 1 : 0;  // the transform has the negative sign
        rc = run_interp(p, c, (int)nrun, st);
        p->stage_half = 0;
        p->stage_conj = 0;
        return rc;
      }
    }
    {
      ScopedTiming tm(kTDeconvPad, st);
      const int64_t nrows = p->ntrans * p->nf[0] * p->nf[1];
      hipLaunchKernelGGL((k_deconv_pad_r<T>), dim3((unsigned)ceil_div(nrows, kBlock / 64)), dim3(kBlock), 0, st,
                         (const T*)f, (T*)p->rbuf, p->ntrans, nrun,
                         (int)p->N[0], (int)p->N[1], (int)p->N[2], (int)p->nf[0], (int)p->nf[1], (int)p->nf[2],
                         (const T*)p->invp[0], (const T*)p->invp[1], (const T*)p->invp[2]);
      COV3D_LAUNCH_CHECK();
    }
    rc = exec_fft_real(p, p->fft_r2c, p->rbuf, p->fine, st);
    if (rc) return rc;
    p->stage_half = (int)(p->nf[2] / 2 + 1);
    p->stage_conj = p->iflag2 > 0 ? 1 : 0;  // R2C has the negative sign
    rc = run_interp(p, c, (int)nrun, st);
    p->stage_half = 0;
    p->stage_conj = 0;
    return rc;
  }
  {
    ScopedTiming tm(kTDeconvPad, st);
    if (f_real)
      hipLaunchKernelGGL((k_deconv_pad<T, true>), g1, dim3(kBlock), 0, st, (const T*)f,
                         (Cplx<T>*)p->fine, p->ntrans, nrun, (int)p->N[0], (int)p->N[1],
                         (int)p->N[2], (int)p->nf[0], (int)p->nf[1], (int)p->nf[2],
                         (const T*)p->invp[0], (const T*)p->invp[1], (const T*)p->invp[2]);
    else
      hipLaunchKernelGGL((k_deconv_pad<T, false>), g1, dim3(kBlock), 0, st, (const T*)f,
                         (Cplx<T>*)p->fine, p->ntrans, nrun, (int)p->N[0], (int)p->N[1],
                         (int)p->N[2], (int)p->nf[0], (int)p->nf[1], (int)p->nf[2],
                         (const T*)p->invp[0], (const T*)p->invp[1], (const T*)p->invp[2]);
  }
  COV3D_LAUNCH_CHECK();
  int rc = exec_fft(p, p->iflag2 < 0, st);
  if (rc) return rc;
  return run_interp(p, c, (int)nrun, st);
}

template <typename T>
static int exec_type1(Plan* p, const void* c, void* f, bool f_real, int64_t nrun, hipStream_t st) {
  if (p->M == 0) {
    // no points: the spread grid is all zeros
    const int64_t total = p->nf[0] * p->nf[1] * p->nf[2] * p->ntrans * 2;
    hipLaunchKernelGGL((k_fill_zero<T>), dim3((unsigned)std::min<int64_t>(ceil_div(total, kBlock), 65536)),
                       dim3(kBlock), 0, st, (T*)p->fine, total);
    COV3D_LAUNCH_CHECK();
  } else {
    int rc = run_spread<T>(p, c, (int)nrun, st);
    if (rc) return rc;
  }
  const int64_t nm = p->N[0] * p->N[1] * p->N[2];
  dim3 g1((unsigned)ceil_div(nm * nrun, kBlock));
  if (f_real && p->real_fft) {
    // real output: Hermitian half of the spread grid -> C2R (positive sign; the half is conjugated
    // when type 1 takes the negative sign) -> real fine grid -> deconvolved real modes
    int rc = ensure_real(p);
    if (rc) return rc;
    if constexpr (sizeof(T) == 4) {
      if (p->pruned) {
        ScopedTiming tm(kTFft, st);
        return pruned_type1(p->fine, (float*)f, p->rbuf, p->tw, (const float*)p->invp[0], (const float*)p->invp[1],
                            (const float*)p->invp[2], (int)nrun, (int)p->nf[0], (int)p->N[0],
                            p->iflag2 > 0 ? 1 : 0, st);
      }
    }
    const int64_t nh = p->nf[0] * p->nf[1] * (p->nf[2] / 2 + 1) * p->ntrans;
    {
      ScopedTiming tm(kTDeconvCrop, st);
      hipLaunchKernelGGL((k_fold_half<T>), dim3((unsigned)ceil_div(nh, kBlock)), dim3(kBlock), 0, st,
                         (const Cplx<T>*)p->fine, (Cplx<T>*)p->rbuf, p->ntrans, (int)p->nf[0], (int)p->nf[1],
                         (int)p->nf[2], p->iflag2 > 0 ? 1 : 0);
      COV3D_LAUNCH_CHECK();
    }
    rc = exec_fft_real(p, p->fft_c2r, p->rbuf, p->fine, st);
    if (rc) return rc;
    ScopedTiming tm(kTDeconvCrop, st);
    hipLaunchKernelGGL((k_deconv_crop_r<T>), g1, dim3(kBlock), 0, st, (const T*)p->fine, (T*)f, nrun,
                       (int)p->N[0], (int)p->N[1], (int)p->N[2], (int)p->nf[0], (int)p->nf[1], (int)p->nf[2],
                       (const T*)p->invp[0], (const T*)p->invp[1], (const T*)p->invp[2]);
    COV3D_LAUNCH_CHECK();
    return 0;
  }
  // type 1 uses the opposite sign of type 2
  int rc = exec_fft(p, !(p->iflag2 < 0), st);
  if (rc) return rc;
  ScopedTiming tm(kTDeconvCrop, st);
  if (f_real)
    hipLaunchKernelGGL((k_deconv_crop<T, true>), g1, dim3(kBlock), 0, st, (const Cplx<T>*)p->fine,
                       (T*)f, nrun, (int)p->N[0], (int)p->N[1], (int)p->N[2], (int)p->nf[0],
                       (int)p->nf[1], (int)p->nf[2], (const T*)p->invp[0], (const T*)p->invp[1],
                       (const T*)p->invp[2]);
  else
    hipLaunchKernelGGL((k_deconv_crop<T, false>), g1, dim3(kBlock), 0, st, (const Cplx<T>*)p->fine,
                       (T*)f, nrun, (int)p->N[0], (int)p->N[1], (int)p->N[2], (int)p->nf[0],
                       (int)p->nf[1], (int)p->nf[2], (const T*)p->invp[0], (const T*)p->invp[1],
                       (const T*)p->invp[2]);
  COV3D_LAUNCH_CHECK();
  return 0;
}

template <typename T>
static int setpts_impl(Plan* p, int64_t M, const T* x, const T* y, const T* z, hipStream_t st) {
  // subproblems: at most one per non-empty super-bin plus one per kMaxSub points
  const int64_t cap = ceil_div(std::min<int64_t>(p->nsuper, M) + ceil_div(M, kMaxSub), 8 * kSubChunk) * 8 * kSubChunk;
  if (M > p->M_cap || cap > p->sub_alloc) {
    free_points(p);
    const int64_t Ma = M > 0 ? M : 1;
    COV3D_HIP(hipMalloc(&p->pt_bin, Ma * sizeof(int)));
    COV3D_HIP(hipMalloc(&p->pt_pos, Ma * sizeof(int)));
    COV3D_HIP(hipMalloc(&p->sidx, Ma * sizeof(int)));
    COV3D_HIP(hipMalloc(&p->spos, Ma * sizeof(int)));
    COV3D_HIP(hipMalloc(&p->irec, Ma * sizeof(float4)));
    COV3D_HIP(hipMalloc(&p->iidx, Ma * sizeof(int)));
    // interpolation writes whole chunks of up to 4 transforms (ntrans rounded up to 4); the
    // spreading input uses row stride gather_ldc(n) for any n <= ntrans.  The matrix-core
    // spreading reads up to 7 values past a row's chunk: a zeroed tail keeps the last row's
    // reads inside the buffer.
    {
      const size_t csz = p->dtype == 1 ? 16 : 8;
      const size_t body = (size_t)Ma * p->ldc_cap * csz;
      COV3D_HIP(hipMalloc(&p->csort, body + kCsortSlack * csz));
      COV3D_HIP(hipMemset((char*)p->csort + body, 0, kCsortSlack * csz));
    }
    for (int k = 0; k < 3; ++k) COV3D_HIP(hipMalloc(&p->sv[k], Ma * sizeof(T)));
    COV3D_HIP(hipMalloc(&p->pv, Ma * 4 * sizeof(T)));
    const int64_t ca = cap > 8 ? cap : 8;
    COV3D_HIP(hipMalloc(&p->sub_sb, ca * sizeof(int)));
    COV3D_HIP(hipMalloc(&p->sub_pt, ca * sizeof(int)));
    COV3D_HIP(hipMalloc(&p->sub_n, ca * sizeof(int)));
    p->M_cap = M;
    p->sub_alloc = ca;
  }
  p->M = M;
  p->sub_cap = cap;
  ScopedTiming tm(kTSort, st);
  hipLaunchKernelGGL(k_fill_int, dim3((unsigned)ceil_div(p->nbins, kBlock)), dim3(kBlock), 0, st,
                     p->bin_count, p->nbins, 0);
  COV3D_LAUNCH_CHECK();
  if (M > 0) {
    dim3 gm((unsigned)ceil_div(M, kBlock));
    hipLaunchKernelGGL((k_bin_count<T>), gm, dim3(kBlock), 0, st, x, y, z, M, p->geom, p->bin_count,
                       p->pt_bin, (T*)p->pv, p->w);
    COV3D_LAUNCH_CHECK();
    int nbits = 1;
    while ((int64_t)1 << nbits < p->nbins) ++nbits;
    nbits += 3;  // k_bin_count's key: bin * 8 + axis-0 stencil position
    size_t need = 0;
    COV3D_HIP(sort_points_by_bin(nullptr, need, (const unsigned*)p->pt_bin, (unsigned*)p->pt_pos, p->sidx,
                                 (int)M, nbits, st));
    if (need > p->sort_tmp_bytes) {
      COV3D_HIP(hipStreamSynchronize(st));
      hipFree(p->sort_tmp);
      p->sort_tmp = nullptr;
      p->sort_tmp_bytes = 0;
      COV3D_HIP(hipMalloc(&p->sort_tmp, need));
      p->sort_tmp_bytes = need;
    }
    COV3D_HIP(sort_points_by_bin(p->sort_tmp, need, (const unsigned*)p->pt_bin, (unsigned*)p->pt_pos, p->sidx,
                                 (int)M, nbits, st));
  }
  {
    // bin starts and subproblem starts: device-wide exclusive scans
    size_t need1 = 0, need2 = 0;
    COV3D_HIP(exclusive_scan_ints(nullptr, need1, p->bin_count, p->bin_start, (int)p->nbins, st));
    COV3D_HIP(exclusive_scan_ints(nullptr, need2, p->sub_cnt, p->sub_start, (int)p->nsuper, st));
    const size_t need = std::max(need1, need2);
    if (need > p->sort_tmp_bytes) {
      COV3D_HIP(hipStreamSynchronize(st));
      hipFree(p->sort_tmp);
      p->sort_tmp = nullptr;
      p->sort_tmp_bytes = 0;
      COV3D_HIP(hipMalloc(&p->sort_tmp, need));
      p->sort_tmp_bytes = need;
    }
    COV3D_HIP(exclusive_scan_ints(p->sort_tmp, need1, p->bin_count, p->bin_start, (int)p->nbins, st));
    hipLaunchKernelGGL(k_sub_count, dim3((unsigned)ceil_div(p->nsuper, kBlock)), dim3(kBlock), 0, st,
                       p->bin_count, p->nsuper, p->sub_cnt);
    COV3D_LAUNCH_CHECK();
    COV3D_HIP(exclusive_scan_ints(p->sort_tmp, need2, p->sub_cnt, p->sub_start, (int)p->nsuper, st));
    hipLaunchKernelGGL(k_sub_total, dim3(1), dim3(64), 0, st, p->sub_cnt, p->sub_start, p->nsuper, p->nsub);
    COV3D_LAUNCH_CHECK();
  }
  if (M > 0) {
    hipLaunchKernelGGL(k_sub_fill, dim3((unsigned)ceil_div(p->nsuper, kBlock)), dim3(kBlock), 0, st,
                       p->bin_count, p->bin_start, p->sub_start, p->nsuper, p->sub_sb, p->sub_pt,
                       p->sub_n);
    COV3D_LAUNCH_CHECK();
    dim3 gm((unsigned)ceil_div(M, kBlock));
    hipLaunchKernelGGL((k_bin_place<T>), gm, dim3(kBlock), 0, st, (const T*)p->pv, M, p->sidx, p->spos,
                       (T*)p->sv[0], (T*)p->sv[1], (T*)p->sv[2]);
    COV3D_LAUNCH_CHECK();
  }
  p->irec_valid = false;
  if constexpr (sizeof(T) == 4) {
    if (M > 0 && p->tc == 2 && p->w <= 8) {
      SubDesc sd{p->sub_sb, p->sub_pt, p->sub_n, p->nsub};
      const float* s0 = (const float*)p->sv[0];
      const float* s1 = (const float*)p->sv[1];
      const float* s2 = (const float*)p->sv[2];
      const dim3 gsub((unsigned)p->sub_cap);
      switch (p->w) {
#define COV3D_W(WW)                                                                                      \
  case WW:                                                                                               \
    hipLaunchKernelGGL((k_bank_order<WW>), gsub, dim3(kBlock), 0, st, sd, s0, s1, s2, p->sidx, p->irec, \
                       p->iidx);                                                                          \
    break;
        COV3D_W(2) COV3D_W(3) COV3D_W(4) COV3D_W(5) COV3D_W(6) COV3D_W(7) COV3D_W(8)
#undef COV3D_W
        default: break;
      }
      COV3D_LAUNCH_CHECK();
      p->irec_valid = true;
    }
  }
  p->pts_set = true;
  return 0;
}

static int destroy_plan(Plan* p) {
  if (!p) return 0;
  free_points(p);
  hipFree(p->fine);
  hipFree((void*)p->geom.xcd_tiles);
  hipFree(p->pipe_ctr);
  for (int k = 0; k < 3; ++k) hipFree(p->invp[k]);
  hipFree(p->bin_count);
  hipFree(p->bin_start);
  hipFree(p->sub_start);
  hipFree(p->sub_cnt);
  hipFree(p->nsub);
  hipFree(p->fft_work);
  hipFree(p->rbuf);
  hipFree(p->tw);
  if (p->fft_fwd) rocfft_plan_destroy(p->fft_fwd);
  if (p->fft_bwd) rocfft_plan_destroy(p->fft_bwd);
  if (p->fft_r2c) rocfft_plan_destroy(p->fft_r2c);
  if (p->fft_c2r) rocfft_plan_destroy(p->fft_c2r);
  if (p->info) rocfft_execution_info_destroy(p->info);
  delete p;
  return 0;
}

// Smallest integer >= n that is a multiple of 8 with only 2, 3, 5 as prime factors.
static int64_t next_fine_size(int64_t n) {
  int64_t m = next235even(n);
  while (m % kSuperEdge) m = next235even(m + 2);
  return m;
}

static int make_plan(int type, const int64_t* n, int iflag, int64_t ntrans, double eps, int dtype,
                     Plan** out) {
  Plan* p = new Plan();
  p->type = type;
  p->dtype = dtype;
  p->iflag2 = iflag == 0 ? -1 : (type == 1 ? -iflag : iflag);
  p->ntrans = ntrans;
  double emin = dtype == 0 ? 10 * 1.1920929e-07 : 10 * 2.220446049250313e-16;
  p->eps = eps < emin ? emin : eps;
  p->w = es_width(p->eps);
  p->beta = es_beta(p->w);
  p->poly = es_poly(p->w, p->beta);
  // the ES kernel's exp/sqrt form is the default; COV3D_ES_POLY=1 selects the polynomial form
  // in FP32 interpolation (measured 8.58 -> 8.38 ms at L=128, r=20, 1024 images)
  p->poly.on = env_int("COV3D_ES_POLY", 0) != 0 ? 1 : 0;
  p->tc = dtype == 1 ? 1 : 2;  // interpolation transforms per tile: two FP32 values make one 16-byte cell
  // the persistent pipelined interpolation (default) or the one-workgroup-per-item kernel with the
  // same arithmetic (COV3D_INTERP_PIPE=0; the tests compare the two bit for bit)
  p->interp_pipe = env_int("COV3D_INTERP_PIPE", 1) != 0 ? 1 : 0;
  p->max_tc_gather = 20;
  p->ldc_cap = (int)ceil_div(ntrans, 4) * 4;
  for (int n_run = 1; n_run <= ntrans; ++n_run)
    p->ldc_cap = std::max(p->ldc_cap, gather_ldc(n_run, p->max_tc_gather));
  const size_t csize = dtype == 1 ? 16 : 8;
  for (int k = 0; k < 3; ++k) {
    p->N[k] = n[k];
    p->nf[k] = next_fine_size(std::max<int64_t>(2 * n[k], 2 * p->w));
    p->geom.nf[k] = (int)p->nf[k];
    p->geom.nb[k] = (int)(p->nf[k] / kBinEdge);
    p->geom.nsb[k] = (int)(p->nf[k] / kSuperEdge);
  }
  p->nsuper = (int64_t)p->geom.nsb[0] * p->geom.nsb[1] * p->geom.nsb[2];
  p->nbins = 8 * p->nsuper;
  p->geom.xcd_tiles = nullptr;
  const int64_t nfine = p->nf[0] * p->nf[1] * p->nf[2];
  auto fail = [&](int rc) {
    destroy_plan(p);
    return rc;
  };
  {
    // spreading's per-XCD tile order (gather_bin): tiles by wrapped distance from frequency 0, the
    // fine grid's corner, where the slices' points are densest; nearest first
    constexpr int TB = 8;
    const int nb0 = p->geom.nb[0], nb1 = p->geom.nb[1], nb2 = p->geom.nb[2];
    if (nb0 % TB == 0 && nb1 % TB == 0 && nb2 % TB == 0 && (nb0 / TB) % 8 == 0) {
      const int T0 = nb0 / TB, T1 = nb1 / TB, T2 = nb2 / TB;
      const int per = T0 * T1 * T2 / 8;
      std::vector<int> tab((size_t)8 * per);
      auto wrapped = [](int t, int T) {  // twice the distance of tile t's centre from 0 (mod T)
        const int c = 2 * t + 1;
        return std::min(c, 2 * T - c);
      };
      for (int x = 0; x < 8; ++x) {
        std::vector<std::pair<int64_t, int>> key((size_t)per);
        for (int j = 0; j < per; ++j) {
          const int m = j / (T1 * T2), rr = j % (T1 * T2), t1 = rr / T2, t2 = rr % T2;
          const int t0 = ((x - 2 * t1 - 4 * t2) & 7) + 8 * m;
          const int64_t d0 = wrapped(t0, T0), d1 = wrapped(t1, T1), d2 = wrapped(t2, T2);
          key[j] = {d0 * d0 + d1 * d1 + d2 * d2, j};
        }
        std::stable_sort(key.begin(), key.end(),
                         [](const std::pair<int64_t, int>& a, const std::pair<int64_t, int>& b) { return a.first < b.first; });
        for (int j = 0; j < per; ++j) tab[(size_t)x * per + j] = key[j].second;
      }
      int* d = nullptr;
      if (hipMalloc(&d, tab.size() * sizeof(int)) != hipSuccess) return fail(COV3D_ERR_HIP);
      p->geom.xcd_tiles = d;
      if (hipMemcpy(d, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
        return fail(COV3D_ERR_HIP);
    }
  }
  if (hipMalloc(&p->fine, nfine * ntrans * csize) != hipSuccess) {
    set_error("hipMalloc of %lld-byte fine grid failed", (long long)(nfine * ntrans * csize));
    return fail(COV3D_ERR_HIP);
  }
  for (int k = 0; k < 3; ++k) {
    std::vector<double> ip = deconv_factors(p->N[k], p->nf[k], p->w, p->beta);
    if (hipMalloc(&p->invp[k], p->N[k] * (dtype ? 8 : 4)) != hipSuccess) return fail(COV3D_ERR_HIP);
    if (dtype == 1) {
      if (hipMemcpy(p->invp[k], ip.data(), p->N[k] * 8, hipMemcpyHostToDevice) != hipSuccess)
        return fail(COV3D_ERR_HIP);
    } else {
      std::vector<float> ipf(ip.begin(), ip.end());
      if (hipMemcpy(p->invp[k], ipf.data(), p->N[k] * 4, hipMemcpyHostToDevice) != hipSuccess)
        return fail(COV3D_ERR_HIP);
    }
  }
  if (hipMalloc(&p->bin_count, p->nbins * sizeof(int)) != hipSuccess ||
      hipMalloc(&p->bin_start, p->nbins * sizeof(int)) != hipSuccess ||
      hipMalloc(&p->sub_start, p->nsuper * sizeof(int)) != hipSuccess ||
      hipMalloc(&p->sub_cnt, p->nsuper * sizeof(int)) != hipSuccess ||
      hipMalloc(&p->nsub, sizeof(int)) != hipSuccess) {
    set_error("hipMalloc of bin tables failed");
    return fail(COV3D_ERR_HIP);
  }
  std::call_once(g_rocfft_once, []() { rocfft_setup(); });
  size_t lengths[3] = {(size_t)p->nf[2], (size_t)p->nf[1], (size_t)p->nf[0]};
  rocfft_precision prec = dtype == 1 ? rocfft_precision_double : rocfft_precision_single;
  if (rocfft_plan_create(&p->fft_fwd, rocfft_placement_inplace, rocfft_transform_type_complex_forward,
                         prec, 3, lengths, (size_t)ntrans, nullptr) != rocfft_status_success ||
      rocfft_plan_create(&p->fft_bwd, rocfft_placement_inplace,
                         rocfft_transform_type_complex_inverse, prec, 3, lengths, (size_t)ntrans,
                         nullptr) != rocfft_status_success ||
      rocfft_execution_info_create(&p->info) != rocfft_status_success) {
    set_error("rocfft plan creation failed for %lldx%lldx%lld x %lld", (long long)p->nf[0],
              (long long)p->nf[1], (long long)p->nf[2], (long long)ntrans);
    return fail(COV3D_ERR_FFT);
  }
  size_t w1 = 0, w2 = 0;
  rocfft_plan_get_work_buffer_size(p->fft_fwd, &w1);
  rocfft_plan_get_work_buffer_size(p->fft_bwd, &w2);
  p->real_fft = 1;
  {
    const size_t tsz = dtype == 1 ? 8 : 4;
    p->rbuf_bytes = (size_t)ntrans * std::max<size_t>((size_t)nfine * tsz,
                                                        (size_t)p->nf[0] * p->nf[1] * (p->nf[2] / 2 + 1) * 2 * tsz);
  }
  if (dtype == 0 && env_int("COV3D_FFT_MODE", 1) == 1 && p->nf[0] == p->nf[1] && p->nf[0] == p->nf[2] &&
      p->N[0] == p->N[1] && p->N[0] == p->N[2] && pruned_fft_supported(p->nf[0], p->N[0]) &&
      (size_t)pruned_fft_work(p->nf[0], p->N[0], ntrans) * 8 <= p->rbuf_bytes) {
    const int64_t NF = p->nf[0];
    std::vector<float> tw(2 * NF);
    for (int64_t k = 0; k < NF; ++k) {
      tw[2 * k] = (float)cos(2.0 * M_PI * (double)k / (double)NF);
      tw[2 * k + 1] = (float)-sin(2.0 * M_PI * (double)k / (double)NF);
    }
    if (hipMalloc(&p->tw, NF * 8) != hipSuccess) return fail(COV3D_ERR_HIP);
    if (hipMemcpy(p->tw, tw.data(), NF * 8, hipMemcpyHostToDevice) != hipSuccess) return fail(COV3D_ERR_HIP);
    p->pruned = true;
  }
  p->work_size = std::max(w1, w2);
  if (p->work_size) {
    if (hipMalloc(&p->fft_work, p->work_size) != hipSuccess) return fail(COV3D_ERR_HIP);
    if (rocfft_execution_info_set_work_buffer(p->info, p->fft_work, p->work_size) !=
        rocfft_status_success)
      return fail(COV3D_ERR_FFT);
  }
  *out = p;
  return 0;
}

}  // namespace cov3d

using cov3d::Plan;

extern "C" int cov3d_nufft_makeplan(int type, int64_t n0, int64_t n1, int64_t n2, int iflag,
                                    int64_t ntrans, double eps, int dtype, void** plan) {
  COV3D_CHECK_ARG(plan != nullptr, "plan pointer is NULL");
  COV3D_CHECK_ARG(type >= 0 && type <= 2, "type must be 0, 1 or 2 (got %d)", type);
  COV3D_CHECK_ARG(n0 >= 1 && n1 >= 1 && n2 >= 1, "mode counts must be >= 1");
  COV3D_CHECK_ARG(ntrans >= 1, "ntrans must be >= 1");
  COV3D_CHECK_ARG(dtype == 0 || dtype == 1, "dtype must be 0 (f32) or 1 (f64)");
  COV3D_CHECK_ARG(eps > 0, "eps must be positive");
  int64_t n[3] = {n0, n1, n2};
  Plan* p = nullptr;
  int rc = cov3d::make_plan(type, n, iflag, ntrans, eps, dtype, &p);
  if (rc) return rc;
  *plan = p;
  return 0;
}

extern "C" int cov3d_nufft_setpts(void* plan, int64_t M, const void* x, const void* y,
                                  const void* z, void* stream) {
  COV3D_CHECK_ARG(plan != nullptr, "plan is NULL");
  COV3D_CHECK_ARG(M >= 0 && M < (int64_t)INT32_MAX, "point count %lld out of range", (long long)M);
  COV3D_CHECK_ARG(M == 0 || (x && y && z), "point arrays must not be NULL");
  Plan* p = static_cast<Plan*>(plan);
  hipStream_t st = cov3d::as_stream(stream);
  if (p->dtype == 1)
    return cov3d::setpts_impl<double>(p, M, (const double*)x, (const double*)y, (const double*)z, st);
  return cov3d::setpts_impl<float>(p, M, (const float*)x, (const float*)y, (const float*)z, st);
}

extern "C" int cov3d_nufft_execute_dir(void* plan, int type, void* c, void* f, int f_is_real,
                                       int64_t ntrans_run, void* stream) {
  COV3D_CHECK_ARG(plan != nullptr, "plan is NULL");
  Plan* p = static_cast<Plan*>(plan);
  COV3D_CHECK_ARG(type == 1 || type == 2, "direction must be 1 or 2");
  COV3D_CHECK_ARG(p->type == 0 || p->type == type, "plan of type %d cannot run type %d", p->type,
                  type);
  if (!p->pts_set) {
    cov3d::set_error("execute called before setpts");
    return COV3D_ERR_STATE;
  }
  if (ntrans_run <= 0) ntrans_run = p->ntrans;
  COV3D_CHECK_ARG(ntrans_run <= p->ntrans, "ntrans_run %lld > plan ntrans %lld",
                  (long long)ntrans_run, (long long)p->ntrans);
  COV3D_CHECK_ARG(f != nullptr && (c != nullptr || p->M == 0), "data pointers must not be NULL");
  hipStream_t st = cov3d::as_stream(stream);
  if (type == 2) {
    if (p->dtype == 1) return cov3d::exec_type2<double>(p, f, f_is_real != 0, c, ntrans_run, st);
    return cov3d::exec_type2<float>(p, f, f_is_real != 0, c, ntrans_run, st);
  }
  if (p->dtype == 1) return cov3d::exec_type1<double>(p, c, f, f_is_real != 0, ntrans_run, st);
  return cov3d::exec_type1<float>(p, c, f, f_is_real != 0, ntrans_run, st);
}

extern "C" int cov3d_nufft_sorted_input(void* plan, int64_t ntrans_run, void** values, int64_t* ldc,
                                        void** spos) {
  COV3D_CHECK_ARG(plan != nullptr && values && ldc && spos, "NULL pointer");
  Plan* p = static_cast<Plan*>(plan);
  COV3D_CHECK_ARG(p->type == 0 || p->type == 1, "plan of type %d has no type-1 input", p->type);
  if (!p->pts_set) {
    cov3d::set_error("sorted_input called before setpts");
    return COV3D_ERR_STATE;
  }
  if (ntrans_run <= 0) ntrans_run = p->ntrans;
  COV3D_CHECK_ARG(ntrans_run <= p->ntrans, "ntrans_run %lld > plan ntrans %lld", (long long)ntrans_run,
                  (long long)p->ntrans);
  *values = p->csort;
  *ldc = cov3d::spread_ldc(p, (int)ntrans_run);
  *spos = p->spos;
  return 0;
}

extern "C" int cov3d_nufft_adjoint_presorted(void* plan, void* f, int f_is_real, int64_t ntrans_run,
                                             void* stream) {
  COV3D_CHECK_ARG(plan != nullptr, "plan is NULL");
  Plan* p = static_cast<Plan*>(plan);
  COV3D_CHECK_ARG(p->type == 0 || p->type == 1, "plan of type %d cannot run type 1", p->type);
  if (!p->pts_set) {
    cov3d::set_error("execute called before setpts");
    return COV3D_ERR_STATE;
  }
  if (ntrans_run <= 0) ntrans_run = p->ntrans;
  COV3D_CHECK_ARG(ntrans_run <= p->ntrans, "ntrans_run %lld > plan ntrans %lld", (long long)ntrans_run,
                  (long long)p->ntrans);
  COV3D_CHECK_ARG(f != nullptr, "output pointer must not be NULL");
  hipStream_t st = cov3d::as_stream(stream);
  p->presorted = true;
  const int rc = p->dtype == 1 ? cov3d::exec_type1<double>(p, p->csort, f, f_is_real != 0, ntrans_run, st)
                               : cov3d::exec_type1<float>(p, p->csort, f, f_is_real != 0, ntrans_run, st);
  p->presorted = false;
  return rc;
}

extern "C" int cov3d_nufft_execute(void* plan, void* c, void* f, int f_is_real, int64_t ntrans_run,
                                   void* stream) {
  COV3D_CHECK_ARG(plan != nullptr, "plan is NULL");
  Plan* p = static_cast<Plan*>(plan);
  COV3D_CHECK_ARG(p->type != 0, "type-0 plans need cov3d_nufft_execute_dir");
  return cov3d_nufft_execute_dir(plan, p->type, c, f, f_is_real, ntrans_run, stream);
}

extern "C" int cov3d_nufft_destroy(void* plan) {
  return cov3d::destroy_plan(static_cast<Plan*>(plan));
}

extern "C" int cov3d_nufft_info(void* plan, int64_t* host_nf3, int* host_w, double* host_beta) {
  COV3D_CHECK_ARG(plan != nullptr, "plan is NULL");
  Plan* p = static_cast<Plan*>(plan);
  if (host_nf3)
    for (int k = 0; k < 3; ++k) host_nf3[k] = p->nf[k];
  if (host_w) *host_w = p->w;
  if (host_beta) *host_beta = p->beta;
  return 0;
}

// Enables (1) or disables (0) event timing of the NUFFT kernel classes and
// drops every recorded interval.
extern "C" int cov3d_timing_enable(int on) {
  for (int k = 0; k < cov3d::kTCount; ++k) {
    for (auto& e : cov3d::g_timing.ev[k]) {
      hipEventDestroy(e.first);
      hipEventDestroy(e.second);
    }
    cov3d::g_timing.ev[k].clear();
  }
  cov3d::g_timing.enabled = on != 0;
  return 0;
}

// Sums the recorded intervals of one kernel class (0 interp, 1 spread,
// 2 deconv/pad, 3 deconv/crop, 4 rocFFT, 5 bin sort, 6 point permutes); waits for them.
extern "C" int cov3d_timing_read(int kind, double* host_total_ms, int64_t* host_count) {
  COV3D_CHECK_ARG(kind >= 0 && kind < cov3d::kTCount, "bad timing kind %d", kind);
  double tot = 0;
  for (auto& e : cov3d::g_timing.ev[kind]) {
    COV3D_HIP(hipEventSynchronize(e.second));
    float ms = 0;
    COV3D_HIP(hipEventElapsedTime(&ms, e.first, e.second));
    tot += ms;
  }
  if (host_total_ms) *host_total_ms = tot;
  if (host_count) *host_count = (int64_t)cov3d::g_timing.ev[kind].size();
  return 0;
}
