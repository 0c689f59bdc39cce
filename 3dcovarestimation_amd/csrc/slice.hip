// Discretised Fourier-slice extraction (trilinear / nearest) from an upsampled
// centred-FFT volume and its exact adjoint (scatter-add).
//
// Replaces torch.nn.functional.grid_sample(input=(1, 2r, D, H, W), grid,
// align_corners=True, padding_mode="zeros") in
// NufftPlanDiscretized.execute_forward (cov3d/nufft_plan.py:120-169) and the
// autograd adjoint used by execute_adjoint / execute_adjoint_unaggregated
// (nufft_plan.py:171-242).  The coordinate unnormalisation, corner rounding,
// weight products and the corner order follow grid_sample's documented
// trilinear semantics so the two agree to float rounding.
#include <map>
#include <mutex>

#include "common.h"

namespace cov3d {

constexpr int kSliceBlock = 256;

hipError_t sort_points_by_bin(void* tmp, size_t& tmp_bytes, const unsigned* keys_in, unsigned* keys_out,
                              int* idx_out, int n, int nbits, hipStream_t st);

template <typename T>
struct Corner8 {
  int64_t off[8];
  T w[8];
  bool ok[8];
};

// A point of a slice plan in sorted order: its unnormalised grid coordinates and caller index.
template <typename T>
struct SortedPoint {
  T x, y, z;  // unnormalised grid coordinates
  int j;      // caller index
};

// grid_sample unnormalisation with align_corners=True: ((g + 1) / 2) * (size - 1).
template <typename T>
__device__ __forceinline__ T unnorm(T g, int64_t size) {
  return ((g + (T)1) / (T)2) * (T)(size - 1);
}

// Trilinear corners in the order tnw, tne, tsw, tse, bnw, bne, bsw, bse
// (t/b = z0/z1, n/s = y0/y1, w/e = x0/x1).
template <typename T>
__device__ __forceinline__ void trilinear_corners(T ix, T iy, T iz, int64_t D, int64_t H, int64_t W,
                                                  Corner8<T>& cn) {
  T fx = floor(ix), fy = floor(iy), fz = floor(iz);
  int64_t x0 = (int64_t)fx, y0 = (int64_t)fy, z0 = (int64_t)fz;
  T wx0 = (fx + (T)1) - ix, wx1 = ix - fx;
  T wy0 = (fy + (T)1) - iy, wy1 = iy - fy;
  T wz0 = (fz + (T)1) - iz, wz1 = iz - fz;
  const int64_t xs[8] = {x0, x0 + 1, x0, x0 + 1, x0, x0 + 1, x0, x0 + 1};
  const int64_t ys[8] = {y0, y0, y0 + 1, y0 + 1, y0, y0, y0 + 1, y0 + 1};
  const int64_t zs[8] = {z0, z0, z0, z0, z0 + 1, z0 + 1, z0 + 1, z0 + 1};
  const T wxs[8] = {wx0, wx1, wx0, wx1, wx0, wx1, wx0, wx1};
  const T wys[8] = {wy0, wy0, wy1, wy1, wy0, wy0, wy1, wy1};
  const T wzs[8] = {wz0, wz0, wz0, wz0, wz1, wz1, wz1, wz1};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    cn.w[k] = wxs[k] * wys[k] * wzs[k];
    cn.ok[k] = xs[k] >= 0 && xs[k] < W && ys[k] >= 0 && ys[k] < H && zs[k] >= 0 && zs[k] < D;
    cn.off[k] = cn.ok[k] ? (zs[k] * H + ys[k]) * W + xs[k] : 0;
  }
}

// One thread per (point, transform), the transform the slow grid index: all points of one
// transform run before the next, so the 8 corner reads of consecutive points hit one transform's
// volume (134 MB at vL = 256, resident in the 256 MB Infinity Cache) and every output row is written
// coalesced.  Measured at L = 128, r = 20, 1024 images: 20.8 ms; four transforms per thread (the
// corners formed once, 32 loads in flight) took 42.6 ms, its four-volume working set no longer
// fitting the cache.
constexpr int kFwdGroup = 1;

template <typename T, int MODE>
__global__ void __launch_bounds__(kSliceBlock)
    k_slice_forward(const Cplx<T>* __restrict__ vol, const T* __restrict__ grid,
                    Cplx<T>* __restrict__ out, int64_t D, int64_t H, int64_t W, int64_t npts, int ntrans) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int t0 = blockIdx.y * kFwdGroup;
  if (j >= npts) return;
  const int64_t vsize = D * H * W;
  T ix = unnorm(grid[3 * j], W), iy = unnorm(grid[3 * j + 1], H), iz = unnorm(grid[3 * j + 2], D);
  Cplx<T> acc[kFwdGroup];
#pragma unroll
  for (int u = 0; u < kFwdGroup; ++u) acc[u] = {(T)0, (T)0};
  if (MODE == 0) {
    Corner8<T> cn;
    trilinear_corners(ix, iy, iz, D, H, W, cn);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (cn.ok[k]) {
#pragma unroll
        for (int u = 0; u < kFwdGroup; ++u) {
          if (t0 + u < ntrans) {
            const Cplx<T> a = vol[(int64_t)(t0 + u) * vsize + cn.off[k]];
            acc[u].x += a.x * cn.w[k];
            acc[u].y += a.y * cn.w[k];
          }
        }
      }
    }
  } else {
    int64_t x = (int64_t)rint(ix), y = (int64_t)rint(iy), z = (int64_t)rint(iz);
    if (x >= 0 && x < W && y >= 0 && y < H && z >= 0 && z < D) {
#pragma unroll
      for (int u = 0; u < kFwdGroup; ++u)
        if (t0 + u < ntrans) acc[u] = vol[(int64_t)(t0 + u) * vsize + (z * H + y) * W + x];
    }
  }
#pragma unroll
  for (int u = 0; u < kFwdGroup; ++u)
    if (t0 + u < ntrans) out[(int64_t)(t0 + u) * npts + j] = acc[u];
}

// Gradient of the trilinear slice with respect to its normalised grid coordinates, the
// derivative torch's grid_sample takes for the grid (the pose gradient of the discretised
// projection).  One thread per point, all transforms in a loop: for each transform t the
// output gradient g and the 8 corner values a_k give s_k = Re(conj(g) a_k), and
//   d/d ix = sum_k s_k * (+-1)_x * wy_k * wz_k  (likewise y, z),
// scaled by (size - 1) / 2 from the align_corners unnormalisation.  Corners outside the volume
// contribute nothing (zero padding).  Nearest sampling is piecewise constant: its gradient is 0.
template <typename T, int MODE>
__global__ void __launch_bounds__(kSliceBlock)
    k_slice_point_grad(const Cplx<T>* __restrict__ vol, const T* __restrict__ grid,
                       const Cplx<T>* __restrict__ gout, T* __restrict__ ggrid, int64_t D, int64_t H,
                       int64_t W, int64_t npts, int ntrans) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npts) return;
  T gx = 0, gy = 0, gz = 0;
  if (MODE == 0) {
    const int64_t vsize = D * H * W;
    const T ix = unnorm(grid[3 * j], W), iy = unnorm(grid[3 * j + 1], H), iz = unnorm(grid[3 * j + 2], D);
    Corner8<T> cn;
    trilinear_corners(ix, iy, iz, D, H, W, cn);
    const T fx = floor(ix), fy = floor(iy), fz = floor(iz);
    const T wx[2] = {(fx + (T)1) - ix, ix - fx};
    const T wy[2] = {(fy + (T)1) - iy, iy - fy};
    const T wz[2] = {(fz + (T)1) - iz, iz - fz};
    T s[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] = 0;
    for (int t = 0; t < ntrans; ++t) {
      const Cplx<T> g = gout[(int64_t)t * npts + j];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (cn.ok[k]) {
          const Cplx<T> a = vol[(int64_t)t * vsize + cn.off[k]];
          s[k] += g.x * a.x + g.y * a.y;
        }
      }
    }
    // corner k: x index bit 0, y bit 1, z bit 2 (tnw, tne, tsw, tse, bnw, bne, bsw, bse)
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int bx = k & 1, by = (k >> 1) & 1, bz = (k >> 2) & 1;
      const T sx = bx ? (T)1 : (T)-1, sy = by ? (T)1 : (T)-1, sz = bz ? (T)1 : (T)-1;
      gx += s[k] * sx * wy[by] * wz[bz];
      gy += s[k] * wx[bx] * sy * wz[bz];
      gz += s[k] * wx[bx] * wy[by] * sz;
    }
    gx *= (T)(W - 1) / (T)2;
    gy *= (T)(H - 1) / (T)2;
    gz *= (T)(D - 1) / (T)2;
  }
  ggrid[3 * j] = gx;
  ggrid[3 * j + 1] = gy;
  ggrid[3 * j + 2] = gz;
}

template <typename T, int MODE>
__global__ void __launch_bounds__(kSliceBlock)
    k_slice_adjoint(const Cplx<T>* __restrict__ sig, const T* __restrict__ grid,
                    Cplx<T>* __restrict__ vol, int64_t D, int64_t H, int64_t W, int64_t npts,
                    int aggregated, int64_t pts_per_vol) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t t = blockIdx.y;
  if (j >= npts) return;
  const int64_t vsize = D * H * W;
  const int64_t vidx = aggregated ? t : j / pts_per_vol;
  T* v = reinterpret_cast<T*>(vol + vidx * vsize);
  const Cplx<T> s = sig[t * npts + j];
  T ix = unnorm(grid[3 * j], W), iy = unnorm(grid[3 * j + 1], H), iz = unnorm(grid[3 * j + 2], D);
  if (MODE == 0) {
    Corner8<T> cn;
    trilinear_corners(ix, iy, iz, D, H, W, cn);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (cn.ok[k]) {
        atomicAdd(v + 2 * cn.off[k], s.x * cn.w[k]);
        atomicAdd(v + 2 * cn.off[k] + 1, s.y * cn.w[k]);
      }
    }
  } else {
    int64_t x = (int64_t)rint(ix), y = (int64_t)rint(iy), z = (int64_t)rint(iz);
    if (x >= 0 && x < W && y >= 0 && y < H && z >= 0 && z < D) {
      int64_t off = (z * H + y) * W + x;
      atomicAdd(v + 2 * off, s.x);
      atomicAdd(v + 2 * off + 1, s.y);
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Aggregated adjoint as a gather (no atomics).  Every point has a "floor cell" f = floor(coord)
// (trilinear: it reaches cells f + {0,1}^3) or a rounded cell (nearest: it reaches that cell only).
// The points are radix-sorted by that cell, stored with a +1 offset per axis so that f = -1 is
// representable ((W+1)(H+1)(D+1) keys; points that reach no cell of the volume get the sentinel
// key and sort last), and each point's unnormalised coordinates and r values are written once, in
// the caller's order, to its row at its sorted position.  Output cell q then sums, over the 8
// floor cells q - {0,1}^3 (one cell for nearest), the contiguous runs of rows with that key, each
// times its trilinear weight for q, for all transforms (up to 20 per thread).  The sums run in
// sorted order, so the result is deterministic, and every cell is written once.
// ---------------------------------------------------------------------------------------------

template <typename T, int MODE>
__global__ void __launch_bounds__(kSliceBlock)
    k_slice_keys(const T* __restrict__ grid, unsigned* __restrict__ keys, int64_t D, int64_t H, int64_t W,
                 int64_t npts, unsigned sentinel) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npts) return;
  const T ix = unnorm(grid[3 * j], W), iy = unnorm(grid[3 * j + 1], H), iz = unnorm(grid[3 * j + 2], D);
  int64_t x, y, z;
  bool reach;
  if (MODE == 0) {
    x = (int64_t)floor(ix), y = (int64_t)floor(iy), z = (int64_t)floor(iz);
    reach = x >= -1 && x < W && y >= -1 && y < H && z >= -1 && z < D;
  } else {
    x = (int64_t)rint(ix), y = (int64_t)rint(iy), z = (int64_t)rint(iz);
    reach = x >= 0 && x < W && y >= 0 && y < H && z >= 0 && z < D;
  }
  keys[j] = reach ? (unsigned)(((z + 1) * (H + 1) + (y + 1)) * (W + 1) + (x + 1)) : sentinel;
}

// Run boundaries of the sorted keys: start[k] / end[k] of key k (both zero for absent keys).
__global__ void k_slice_runs(const unsigned* __restrict__ skeys, int* __restrict__ start, int* __restrict__ end,
                             int64_t n, unsigned sentinel) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned k = skeys[i];
  if (k == sentinel) return;
  if (i == 0 || skeys[i - 1] != k) start[k] = (int)i;
  if (i == n - 1 || skeys[i + 1] != k) end[k] = (int)(i + 1);
}

// Sorted position of every point (the inverse of the sort's permutation).
__global__ void k_slice_inverse(const int* __restrict__ sidx, int* __restrict__ spos, int64_t n) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) spos[sidx[s]] = (int)s;
}

// Rows of the sorted points: [ix, iy, iz, 0 | value 0 .. value r-1] (RW = 2 + r float2 words), read in
// the caller's order (coalesced) and written at the point's sorted position.
template <typename T>
__global__ void __launch_bounds__(kSliceBlock)
    k_slice_rows(const Cplx<T>* __restrict__ sig, const T* __restrict__ grid, const int* __restrict__ spos,
                 Cplx<T>* __restrict__ rows, int64_t D, int64_t H, int64_t W, int64_t npts, int ntrans, int rw) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npts) return;
  Cplx<T>* row = rows + (int64_t)spos[j] * rw;
  row[0] = {unnorm(grid[3 * j], W), unnorm(grid[3 * j + 1], H)};
  row[1] = {unnorm(grid[3 * j + 2], D), (T)0};
  for (int t = 0; t < ntrans; ++t) row[2 + t] = sig[(int64_t)t * npts + j];
}

// Rows as k_slice_rows for a slice plan: the coordinates come from the point's sorted record
// (the plan keeps no grid pointer); the values are read in the caller's order (coalesced) and
// the row is written at the point's sorted position.
template <typename T>
__global__ void __launch_bounds__(kSliceBlock)
    k_slice_rows_plan(const Cplx<T>* __restrict__ sig, const SortedPoint<T>* __restrict__ sp,
                      const int* __restrict__ spos, Cplx<T>* __restrict__ rows, int64_t npts, int ntrans, int rw) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npts) return;
  const int64_t s = spos[j];
  const SortedPoint<T> p = sp[s];
  Cplx<T>* row = rows + s * rw;
  row[0] = {p.x, p.y};
  row[1] = {p.z, (T)0};
  for (int t = 0; t < ntrans; ++t) row[2 + t] = sig[(int64_t)t * npts + j];
}

template <typename T>
__global__ void k_zero_vals(T* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (T)0;
}

// Output cell q gathers every transform from the rows of the points whose floor cell is q - d,
// d in {0,1}^3 (nearest: the points rounded to q), in sorted order, and writes each sum once.
template <typename T, int MODE, int RT>
__global__ void __launch_bounds__(kSliceBlock)
    k_slice_adjoint_gather(const Cplx<T>* __restrict__ rows, const int* __restrict__ start,
                           const int* __restrict__ end, Cplx<T>* __restrict__ vol, int64_t D, int64_t H,
                           int64_t W, int ntrans, int rw) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t vsize = D * H * W;
  if (q >= vsize) return;
  const int t0 = blockIdx.y * RT;
  const int64_t qx = q % W, qy = (q / W) % H, qz = q / (W * H);
  T ax[RT], ay[RT];
#pragma unroll
  for (int u = 0; u < RT; ++u) ax[u] = ay[u] = (T)0;
  constexpr int NB = MODE == 0 ? 8 : 1;
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int dx = nb & 1, dy = (nb >> 1) & 1, dz = (nb >> 2) & 1;
    const int64_t key = ((qz - dz + 1) * (H + 1) + (qy - dy + 1)) * (W + 1) + (qx - dx + 1);
    const int b = start[key], e = end[key];
    for (int s = b; s < e; ++s) {
      const Cplx<T>* row = rows + (int64_t)s * rw;
      T w = (T)1;
      if (MODE == 0) {
        const Cplx<T> c0 = row[0], c1 = row[1];
        const T fx = floor(c0.x), fy = floor(c0.y), fz = floor(c1.x);
        // q is corner f + d of this point: weight (f + 1 - coord) for d = 0, (coord - f) for d = 1,
        // the products trilinear_corners forms
        const T wx = dx ? c0.x - fx : (fx + (T)1) - c0.x;
        const T wy = dy ? c0.y - fy : (fy + (T)1) - c0.y;
        const T wz = dz ? c1.x - fz : (fz + (T)1) - c1.x;
        w = wx * wy * wz;
      }
#pragma unroll
      for (int u = 0; u < RT; ++u) {
        if (t0 + u < ntrans) {
          const Cplx<T> v = row[2 + t0 + u];
          ax[u] += w * v.x;
          ay[u] += w * v.y;
        }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < RT; ++u)
    if (t0 + u < ntrans) vol[(int64_t)(t0 + u) * vsize + q] = {ax[u], ay[u]};
}

// Sort workspace of the gather adjoint, per device, grown on demand and kept.
struct AdjWork {
  size_t cap_pts = 0, cap_keys = 0, cap_tmp = 0, cap_rows = 0;
  unsigned *keys = nullptr, *skeys = nullptr;
  int *sidx = nullptr, *spos = nullptr, *start = nullptr, *end = nullptr;
  void* tmp = nullptr;
  void* rows = nullptr;
};
static std::mutex g_adj_mu;
static std::map<int, AdjWork> g_adj;

template <typename T>
__global__ void k_zero_ints(int* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

template <typename T>
static int slice_adj_gather(const void* sig, const void* grid, void* vol, int64_t ntrans, int64_t D, int64_t H,
                            int64_t W, int64_t npts, int mode, hipStream_t st) {
  COV3D_CHECK_ARG(npts < ((int64_t)1 << 31), "too many points for the gather adjoint");
  const int64_t nkeys = (W + 1) * (H + 1) * (D + 1);
  COV3D_CHECK_ARG(nkeys < ((int64_t)1 << 31), "volume too large for the gather adjoint");
  const unsigned sentinel = (unsigned)nkeys;
  int nbits = 1;
  while (((int64_t)1 << nbits) <= nkeys) ++nbits;
  int dev = 0;
  COV3D_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_adj_mu);
  AdjWork& w = g_adj[dev];
  size_t tmp_bytes = 0;
  COV3D_HIP(sort_points_by_bin(nullptr, tmp_bytes, nullptr, nullptr, nullptr, (int)npts, nbits, st));
  if ((size_t)npts > w.cap_pts) {
    hipFree(w.keys), hipFree(w.skeys), hipFree(w.sidx), hipFree(w.spos);
    w.keys = w.skeys = nullptr, w.sidx = w.spos = nullptr, w.cap_pts = 0;
    COV3D_HIP(hipMalloc(&w.keys, npts * sizeof(unsigned)));
    COV3D_HIP(hipMalloc(&w.skeys, npts * sizeof(unsigned)));
    COV3D_HIP(hipMalloc(&w.sidx, npts * sizeof(int)));
    COV3D_HIP(hipMalloc(&w.spos, npts * sizeof(int)));
    w.cap_pts = npts;
  }
  const int rw = 2 + (int)ntrans;
  const size_t row_bytes = (size_t)npts * rw * sizeof(Cplx<T>);
  if (row_bytes > w.cap_rows) {
    hipFree(w.rows);
    w.rows = nullptr, w.cap_rows = 0;
    COV3D_HIP(hipMalloc(&w.rows, row_bytes));
    w.cap_rows = row_bytes;
  }
  if ((size_t)nkeys > w.cap_keys) {
    hipFree(w.start), hipFree(w.end);
    w.start = w.end = nullptr, w.cap_keys = 0;
    COV3D_HIP(hipMalloc(&w.start, nkeys * sizeof(int)));
    COV3D_HIP(hipMalloc(&w.end, nkeys * sizeof(int)));
    w.cap_keys = nkeys;
  }
  if (tmp_bytes > w.cap_tmp) {
    hipFree(w.tmp);
    w.tmp = nullptr, w.cap_tmp = 0;
    COV3D_HIP(hipMalloc(&w.tmp, tmp_bytes));
    w.cap_tmp = tmp_bytes;
  }
  const unsigned nbp = (unsigned)ceil_div(npts, kSliceBlock);
  if (mode == 0)
    hipLaunchKernelGGL((k_slice_keys<T, 0>), dim3(nbp), dim3(kSliceBlock), 0, st, (const T*)grid, w.keys, D, H, W,
                       npts, sentinel);
  else
    hipLaunchKernelGGL((k_slice_keys<T, 1>), dim3(nbp), dim3(kSliceBlock), 0, st, (const T*)grid, w.keys, D, H, W,
                       npts, sentinel);
  COV3D_LAUNCH_CHECK();
  size_t tb = w.cap_tmp;
  COV3D_HIP(sort_points_by_bin(w.tmp, tb, w.keys, w.skeys, w.sidx, (int)npts, nbits, st));
  const unsigned nbk = (unsigned)ceil_div(nkeys, kSliceBlock);
  hipLaunchKernelGGL((k_zero_ints<T>), dim3(nbk), dim3(kSliceBlock), 0, st, w.start, nkeys);
  hipLaunchKernelGGL((k_zero_ints<T>), dim3(nbk), dim3(kSliceBlock), 0, st, w.end, nkeys);
  COV3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_slice_runs, dim3(nbp), dim3(kSliceBlock), 0, st, w.skeys, w.start, w.end, npts, sentinel);
  hipLaunchKernelGGL(k_slice_inverse, dim3(nbp), dim3(kSliceBlock), 0, st, w.sidx, w.spos, npts);
  hipLaunchKernelGGL((k_slice_rows<T>), dim3(nbp), dim3(kSliceBlock), 0, st, (const Cplx<T>*)sig, (const T*)grid,
                     w.spos, (Cplx<T>*)w.rows, D, H, W, npts, (int)ntrans, rw);
  COV3D_LAUNCH_CHECK();
  // all transforms of a cell in one thread up to 20 (the rows stay in the thread's registers once)
  constexpr int RT = 20;
  dim3 g((unsigned)ceil_div(D * H * W, kSliceBlock), (unsigned)ceil_div(ntrans, RT));
  if (mode == 0)
    hipLaunchKernelGGL((k_slice_adjoint_gather<T, 0, RT>), g, dim3(kSliceBlock), 0, st, (const Cplx<T>*)w.rows,
                       w.start, w.end, (Cplx<T>*)vol, D, H, W, (int)ntrans, rw);
  else
    hipLaunchKernelGGL((k_slice_adjoint_gather<T, 1, RT>), g, dim3(kSliceBlock), 0, st, (const Cplx<T>*)w.rows,
                       w.start, w.end, (Cplx<T>*)vol, D, H, W, (int)ntrans, rw);
  COV3D_LAUNCH_CHECK();
  return 0;
}

template <typename T>
static int slice_fwd(const void* vol, const void* grid, void* out, int64_t ntrans, int64_t D,
                     int64_t H, int64_t W, int64_t npts, int mode, hipStream_t st) {
  dim3 g((unsigned)ceil_div(npts, kSliceBlock), (unsigned)ceil_div(ntrans, kFwdGroup));
  if (mode == 0)
    hipLaunchKernelGGL((k_slice_forward<T, 0>), g, dim3(kSliceBlock), 0, st, (const Cplx<T>*)vol,
                       (const T*)grid, (Cplx<T>*)out, D, H, W, npts, (int)ntrans);
  else
    hipLaunchKernelGGL((k_slice_forward<T, 1>), g, dim3(kSliceBlock), 0, st, (const Cplx<T>*)vol,
                       (const T*)grid, (Cplx<T>*)out, D, H, W, npts, (int)ntrans);
  COV3D_LAUNCH_CHECK();
  return 0;
}

template <typename T>
static int slice_pgrad(const void* vol, const void* grid, const void* gout, void* ggrid, int64_t ntrans,
                       int64_t D, int64_t H, int64_t W, int64_t npts, int mode, hipStream_t st) {
  dim3 g((unsigned)ceil_div(npts, kSliceBlock));
  if (mode == 0)
    hipLaunchKernelGGL((k_slice_point_grad<T, 0>), g, dim3(kSliceBlock), 0, st, (const Cplx<T>*)vol,
                       (const T*)grid, (const Cplx<T>*)gout, (T*)ggrid, D, H, W, npts, (int)ntrans);
  else
    hipLaunchKernelGGL((k_slice_point_grad<T, 1>), g, dim3(kSliceBlock), 0, st, (const Cplx<T>*)vol,
                       (const T*)grid, (const Cplx<T>*)gout, (T*)ggrid, D, H, W, npts, (int)ntrans);
  COV3D_LAUNCH_CHECK();
  return 0;
}

template <typename T>
static int slice_adj(const void* sig, const void* grid, void* vol, int64_t ntrans, int64_t D,
                     int64_t H, int64_t W, int64_t npts, int mode, int aggregated, int64_t ppv,
                     hipStream_t st) {
  if (aggregated) return slice_adj_gather<T>(sig, grid, vol, ntrans, D, H, W, npts, mode, st);
  dim3 g((unsigned)ceil_div(npts, kSliceBlock), (unsigned)ntrans);
  if (mode == 0)
    hipLaunchKernelGGL((k_slice_adjoint<T, 0>), g, dim3(kSliceBlock), 0, st, (const Cplx<T>*)sig,
                       (const T*)grid, (Cplx<T>*)vol, D, H, W, npts, aggregated, ppv);
  else
    hipLaunchKernelGGL((k_slice_adjoint<T, 1>), g, dim3(kSliceBlock), 0, st, (const Cplx<T>*)sig,
                       (const T*)grid, (Cplx<T>*)vol, D, H, W, npts, aggregated, ppv);
  COV3D_LAUNCH_CHECK();
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Slice plan: the points sorted once per setpts and shared by the forward and the aggregated
// adjoint.  The forward in caller order reads each point's 8 corners from a volume that does not
// fit in the caches (134 MB per transform at vL = 256), one cache line per corner, and every
// volume cell is read by about 8 points of different images at different times.  In floor-cell
// order the points that share corners are neighbours, so a transform's volume is streamed about
// once and the corner reads of a wavefront hit a few cache lines; each value is then written to
// the point's caller position.
// ---------------------------------------------------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(kSliceBlock)
    k_slice_sorted_points(const T* __restrict__ grid, const int* __restrict__ sidx, SortedPoint<T>* __restrict__ sp,
                          int64_t D, int64_t H, int64_t W, int64_t npts) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= npts) return;
  const int j = sidx[s];
  sp[s] = SortedPoint<T>{unnorm(grid[3 * (int64_t)j], W), unnorm(grid[3 * (int64_t)j + 1], H),
                         unnorm(grid[3 * (int64_t)j + 2], D), j};
}

// The forward of k_slice_forward (same arithmetic, bit for bit) over the points in sorted order,
// one thread per (sorted point, transform), the transform the slow grid index.
template <typename T, int MODE>
__global__ void __launch_bounds__(kSliceBlock)
    k_slice_forward_sorted(const Cplx<T>* __restrict__ vol, const SortedPoint<T>* __restrict__ sp,
                           Cplx<T>* __restrict__ out, int64_t D, int64_t H, int64_t W, int64_t npts) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= npts) return;
  const int64_t t = blockIdx.y;
  const int64_t vsize = D * H * W;
  const Cplx<T>* v = vol + t * vsize;
  const SortedPoint<T> p = sp[s];
  Cplx<T> acc = {(T)0, (T)0};
  if (MODE == 0) {
    Corner8<T> cn;
    trilinear_corners(p.x, p.y, p.z, D, H, W, cn);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (cn.ok[k]) {
        const Cplx<T> a = v[cn.off[k]];
        acc.x += a.x * cn.w[k];
        acc.y += a.y * cn.w[k];
      }
    }
  } else {
    const int64_t x = (int64_t)rint(p.x), y = (int64_t)rint(p.y), z = (int64_t)rint(p.z);
    if (x >= 0 && x < W && y >= 0 && y < H && z >= 0 && z < D) acc = v[(z * H + y) * W + x];
  }
  out[t * npts + p.j] = acc;
}

struct SlicePlan {
  int dev = -1;
  int64_t npts = 0, D = 0, H = 0, W = 0, nkeys = 0;
  int mode = 0, dtype = 0;
  bool ready = false;
  AdjWork w;             // keys, sort, runs, inverse, rows (grown on demand)
  void* sp = nullptr;    // SortedPoint<T>[npts]
  size_t cap_sp = 0;
};

static int plan_grow(void** p, size_t& cap, size_t need) {
  if (need <= cap) return 0;
  hipFree(*p);
  *p = nullptr;
  cap = 0;
  COV3D_HIP(hipMalloc(p, need));
  cap = need;
  return 0;
}

template <typename T>
static int slice_plan_setpts(SlicePlan* pl, const void* grid, int64_t npts, int64_t D, int64_t H, int64_t W,
                             int mode, hipStream_t st) {
  COV3D_CHECK_ARG(npts < ((int64_t)1 << 31), "too many points for a slice plan");
  const int64_t nkeys = (W + 1) * (H + 1) * (D + 1);
  COV3D_CHECK_ARG(nkeys < ((int64_t)1 << 31), "volume too large for a slice plan");
  const unsigned sentinel = (unsigned)nkeys;
  int nbits = 1;
  while (((int64_t)1 << nbits) <= nkeys) ++nbits;
  AdjWork& w = pl->w;
  size_t tmp_bytes = 0;
  COV3D_HIP(sort_points_by_bin(nullptr, tmp_bytes, nullptr, nullptr, nullptr, (int)npts, nbits, st));
  if ((size_t)npts > w.cap_pts) {
    hipFree(w.keys), hipFree(w.skeys), hipFree(w.sidx), hipFree(w.spos);
    w.keys = w.skeys = nullptr, w.sidx = w.spos = nullptr, w.cap_pts = 0;
    COV3D_HIP(hipMalloc(&w.keys, npts * sizeof(unsigned)));
    COV3D_HIP(hipMalloc(&w.skeys, npts * sizeof(unsigned)));
    COV3D_HIP(hipMalloc(&w.sidx, npts * sizeof(int)));
    COV3D_HIP(hipMalloc(&w.spos, npts * sizeof(int)));
    w.cap_pts = npts;
  }
  if ((size_t)nkeys > w.cap_keys) {
    hipFree(w.start), hipFree(w.end);
    w.start = w.end = nullptr, w.cap_keys = 0;
    COV3D_HIP(hipMalloc(&w.start, nkeys * sizeof(int)));
    COV3D_HIP(hipMalloc(&w.end, nkeys * sizeof(int)));
    w.cap_keys = nkeys;
  }
  if (int rc = plan_grow(&w.tmp, w.cap_tmp, tmp_bytes)) return rc;
  if (int rc = plan_grow(&pl->sp, pl->cap_sp, (size_t)npts * sizeof(SortedPoint<T>))) return rc;
  const unsigned nbp = (unsigned)ceil_div(npts, kSliceBlock);
  if (mode == 0)
    hipLaunchKernelGGL((k_slice_keys<T, 0>), dim3(nbp), dim3(kSliceBlock), 0, st, (const T*)grid, w.keys, D, H, W,
                       npts, sentinel);
  else
    hipLaunchKernelGGL((k_slice_keys<T, 1>), dim3(nbp), dim3(kSliceBlock), 0, st, (const T*)grid, w.keys, D, H, W,
                       npts, sentinel);
  COV3D_LAUNCH_CHECK();
  size_t tb = w.cap_tmp;
  COV3D_HIP(sort_points_by_bin(w.tmp, tb, w.keys, w.skeys, w.sidx, (int)npts, nbits, st));
  const unsigned nbk = (unsigned)ceil_div(nkeys, kSliceBlock);
  hipLaunchKernelGGL((k_zero_ints<T>), dim3(nbk), dim3(kSliceBlock), 0, st, w.start, nkeys);
  hipLaunchKernelGGL((k_zero_ints<T>), dim3(nbk), dim3(kSliceBlock), 0, st, w.end, nkeys);
  COV3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_slice_runs, dim3(nbp), dim3(kSliceBlock), 0, st, w.skeys, w.start, w.end, npts, sentinel);
  hipLaunchKernelGGL(k_slice_inverse, dim3(nbp), dim3(kSliceBlock), 0, st, w.sidx, w.spos, npts);
  hipLaunchKernelGGL((k_slice_sorted_points<T>), dim3(nbp), dim3(kSliceBlock), 0, st, (const T*)grid, w.sidx,
                     (SortedPoint<T>*)pl->sp, D, H, W, npts);
  COV3D_LAUNCH_CHECK();
  pl->npts = npts, pl->D = D, pl->H = H, pl->W = W, pl->nkeys = nkeys, pl->mode = mode;
  pl->ready = true;
  return 0;
}

template <typename T>
static int slice_plan_forward(SlicePlan* pl, const void* vol, void* out, int64_t ntrans, hipStream_t st) {
  if (pl->npts == 0) return 0;
  dim3 g((unsigned)ceil_div(pl->npts, kSliceBlock), (unsigned)ntrans);
  if (pl->mode == 0)
    hipLaunchKernelGGL((k_slice_forward_sorted<T, 0>), g, dim3(kSliceBlock), 0, st, (const Cplx<T>*)vol,
                       (const SortedPoint<T>*)pl->sp, (Cplx<T>*)out, pl->D, pl->H, pl->W, pl->npts);
  else
    hipLaunchKernelGGL((k_slice_forward_sorted<T, 1>), g, dim3(kSliceBlock), 0, st, (const Cplx<T>*)vol,
                       (const SortedPoint<T>*)pl->sp, (Cplx<T>*)out, pl->D, pl->H, pl->W, pl->npts);
  COV3D_LAUNCH_CHECK();
  return 0;
}

template <typename T>
static int slice_plan_adjoint(SlicePlan* pl, const void* sig, void* vol, int64_t ntrans, hipStream_t st) {
  AdjWork& w = pl->w;
  const int64_t D = pl->D, H = pl->H, W = pl->W, npts = pl->npts;
  if (npts == 0) {
    const int64_t n = D * H * W * ntrans * 2;
    hipLaunchKernelGGL((k_zero_vals<T>), dim3((unsigned)ceil_div(n, kSliceBlock)), dim3(kSliceBlock), 0, st, (T*)vol, n);
    COV3D_LAUNCH_CHECK();
    return 0;
  }
  const int rw = 2 + (int)ntrans;
  if (int rc = plan_grow(&w.rows, w.cap_rows, (size_t)npts * rw * sizeof(Cplx<T>))) return rc;
  const unsigned nbp = (unsigned)ceil_div(npts, kSliceBlock);
  hipLaunchKernelGGL((k_slice_rows_plan<T>), dim3(nbp), dim3(kSliceBlock), 0, st, (const Cplx<T>*)sig,
                     (const SortedPoint<T>*)pl->sp, w.spos, (Cplx<T>*)w.rows, npts, (int)ntrans, rw);
  COV3D_LAUNCH_CHECK();
  constexpr int RT = 20;
  dim3 g((unsigned)ceil_div(D * H * W, kSliceBlock), (unsigned)ceil_div(ntrans, RT));
  if (pl->mode == 0)
    hipLaunchKernelGGL((k_slice_adjoint_gather<T, 0, RT>), g, dim3(kSliceBlock), 0, st, (const Cplx<T>*)w.rows,
                       w.start, w.end, (Cplx<T>*)vol, D, H, W, (int)ntrans, rw);
  else
    hipLaunchKernelGGL((k_slice_adjoint_gather<T, 1, RT>), g, dim3(kSliceBlock), 0, st, (const Cplx<T>*)w.rows,
                       w.start, w.end, (Cplx<T>*)vol, D, H, W, (int)ntrans, rw);
  COV3D_LAUNCH_CHECK();
  return 0;
}

}  // namespace cov3d

extern "C" int cov3d_slice_forward(const void* vol, const void* grid, void* out, int64_t ntrans,
                                   int64_t D, int64_t H, int64_t W, int64_t npts, int mode,
                                   int dtype, void* stream) {
  COV3D_CHECK_ARG(ntrans >= 1 && D >= 1 && H >= 1 && W >= 1 && npts >= 0, "bad slice shape");
  COV3D_CHECK_ARG(mode == 0 || mode == 1, "mode must be 0 (bilinear) or 1 (nearest)");
  COV3D_CHECK_ARG(dtype == 0 || dtype == 1, "dtype must be 0 or 1");
  COV3D_CHECK_ARG(ntrans < 65536, "ntrans too large");
  if (npts == 0) return 0;
  COV3D_CHECK_ARG(vol && grid && out, "NULL pointer");
  hipStream_t st = cov3d::as_stream(stream);
  if (dtype == 1) return cov3d::slice_fwd<double>(vol, grid, out, ntrans, D, H, W, npts, mode, st);
  return cov3d::slice_fwd<float>(vol, grid, out, ntrans, D, H, W, npts, mode, st);
}

extern "C" int cov3d_slice_adjoint(const void* sig, const void* grid, void* vol_out, int64_t ntrans,
                                   int64_t D, int64_t H, int64_t W, int64_t npts, int mode,
                                   int aggregated, int64_t pts_per_vol, int dtype, void* stream) {
  COV3D_CHECK_ARG(ntrans >= 1 && D >= 1 && H >= 1 && W >= 1 && npts >= 0, "bad slice shape");
  COV3D_CHECK_ARG(mode == 0 || mode == 1, "mode must be 0 (bilinear) or 1 (nearest)");
  COV3D_CHECK_ARG(dtype == 0 || dtype == 1, "dtype must be 0 or 1");
  COV3D_CHECK_ARG(aggregated || (ntrans == 1 && pts_per_vol >= 1),
                  "unaggregated adjoint needs ntrans == 1 and pts_per_vol >= 1");
  COV3D_CHECK_ARG(ntrans < 65536, "ntrans too large");
  if (npts == 0) return 0;
  COV3D_CHECK_ARG(sig && grid && vol_out, "NULL pointer");
  hipStream_t st = cov3d::as_stream(stream);
  if (dtype == 1)
    return cov3d::slice_adj<double>(sig, grid, vol_out, ntrans, D, H, W, npts, mode, aggregated,
                                    pts_per_vol, st);
  return cov3d::slice_adj<float>(sig, grid, vol_out, ntrans, D, H, W, npts, mode, aggregated,
                                 pts_per_vol, st);
}

extern "C" int cov3d_slice_point_grad(const void* vol, const void* grid, const void* grad_out,
                                      void* grid_grad, int64_t ntrans, int64_t D, int64_t H, int64_t W,
                                      int64_t npts, int mode, int dtype, void* stream) {
  COV3D_CHECK_ARG(ntrans >= 1 && D >= 1 && H >= 1 && W >= 1 && npts >= 0, "bad slice shape");
  COV3D_CHECK_ARG(mode == 0 || mode == 1, "mode must be 0 (bilinear) or 1 (nearest)");
  COV3D_CHECK_ARG(dtype == 0 || dtype == 1, "dtype must be 0 or 1");
  if (npts == 0) return 0;
  COV3D_CHECK_ARG(vol && grid && grad_out && grid_grad, "NULL pointer");
  hipStream_t st = cov3d::as_stream(stream);
  if (dtype == 1) return cov3d::slice_pgrad<double>(vol, grid, grad_out, grid_grad, ntrans, D, H, W, npts, mode, st);
  return cov3d::slice_pgrad<float>(vol, grid, grad_out, grid_grad, ntrans, D, H, W, npts, mode, st);
}

extern "C" int cov3d_slice_plan_create(void** plan) {
  COV3D_CHECK_ARG(plan, "NULL plan pointer");
  *plan = new cov3d::SlicePlan();
  return 0;
}

extern "C" int cov3d_slice_plan_setpts(void* plan, const void* grid, int64_t npts, int64_t D, int64_t H, int64_t W,
                                       int mode, int dtype, void* stream) {
  COV3D_CHECK_ARG(plan, "NULL plan");
  COV3D_CHECK_ARG(D >= 1 && H >= 1 && W >= 1 && npts >= 0, "bad slice shape");
  COV3D_CHECK_ARG(mode == 0 || mode == 1, "mode must be 0 (bilinear) or 1 (nearest)");
  COV3D_CHECK_ARG(dtype == 0 || dtype == 1, "dtype must be 0 or 1");
  COV3D_CHECK_ARG(npts == 0 || grid, "NULL grid");
  auto* pl = static_cast<cov3d::SlicePlan*>(plan);
  int dev = 0;
  COV3D_HIP(hipGetDevice(&dev));
  COV3D_CHECK_ARG(pl->dev < 0 || pl->dev == dev, "a slice plan stays on the device it was first used on");
  pl->dev = dev;
  pl->ready = false;
  pl->dtype = dtype;
  if (npts == 0) {
    pl->npts = 0, pl->D = D, pl->H = H, pl->W = W, pl->mode = mode, pl->ready = true;
    return 0;
  }
  hipStream_t st = cov3d::as_stream(stream);
  if (dtype == 1) return cov3d::slice_plan_setpts<double>(pl, grid, npts, D, H, W, mode, st);
  return cov3d::slice_plan_setpts<float>(pl, grid, npts, D, H, W, mode, st);
}

extern "C" int cov3d_slice_plan_forward(void* plan, const void* vol, void* out, int64_t ntrans, void* stream) {
  COV3D_CHECK_ARG(plan, "NULL plan");
  auto* pl = static_cast<cov3d::SlicePlan*>(plan);
  COV3D_CHECK_ARG(pl->ready, "slice plan: setpts first");
  COV3D_CHECK_ARG(ntrans >= 1 && ntrans < 65536, "bad ntrans");
  if (pl->npts == 0) return 0;
  COV3D_CHECK_ARG(vol && out, "NULL pointer");
  hipStream_t st = cov3d::as_stream(stream);
  if (pl->dtype == 1) return cov3d::slice_plan_forward<double>(pl, vol, out, ntrans, st);
  return cov3d::slice_plan_forward<float>(pl, vol, out, ntrans, st);
}

extern "C" int cov3d_slice_plan_adjoint(void* plan, const void* sig, void* vol_out, int64_t ntrans, void* stream) {
  COV3D_CHECK_ARG(plan, "NULL plan");
  auto* pl = static_cast<cov3d::SlicePlan*>(plan);
  COV3D_CHECK_ARG(pl->ready, "slice plan: setpts first");
  COV3D_CHECK_ARG(ntrans >= 1 && ntrans < 65536, "bad ntrans");
  COV3D_CHECK_ARG(vol_out && (sig || pl->npts == 0), "NULL pointer");
  hipStream_t st = cov3d::as_stream(stream);
  if (pl->dtype == 1) return cov3d::slice_plan_adjoint<double>(pl, sig, vol_out, ntrans, st);
  return cov3d::slice_plan_adjoint<float>(pl, sig, vol_out, ntrans, st);
}

extern "C" int cov3d_slice_plan_destroy(void* plan) {
  if (!plan) return 0;
  auto* pl = static_cast<cov3d::SlicePlan*>(plan);
  cov3d::AdjWork& w = pl->w;
  hipFree(w.keys), hipFree(w.skeys), hipFree(w.sidx), hipFree(w.spos), hipFree(w.start), hipFree(w.end);
  hipFree(w.tmp), hipFree(w.rows), hipFree(pl->sp);
  delete pl;
  return 0;
}
