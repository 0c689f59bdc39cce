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
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "common.h"

namespace cov3d {

constexpr int kSliceBlock = 256;

hipError_t sort_points_by_bin(void* tmp, size_t& tmp_bytes, const unsigned* keys_in, unsigned* keys_out,
                              int* idx_out, int n, int nbits, hipStream_t st);
hipError_t exclusive_scan_ints(void* tmp, size_t& tmp_bytes, const int* in, int* out, int n, hipStream_t st);

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
    // a corner outside the volume gets the offset of the nearest cell inside (a valid address
    // near the point's own cells; its value is not used)
    const int64_t xc = xs[k] < 0 ? 0 : (xs[k] >= W ? W - 1 : xs[k]);
    const int64_t yc = ys[k] < 0 ? 0 : (ys[k] >= H ? H - 1 : ys[k]);
    const int64_t zc = zs[k] < 0 ? 0 : (zs[k] >= D ? D - 1 : zs[k]);
    cn.off[k] = (zc * H + yc) * W + xc;
  }
}

// The trilinear sum of one transform's 8 corners (corner order of trilinear_corners; corners outside
// the volume count as zero).  All 8 loads are issued before the first is used: an outside corner
// reads its nearest inside cell and its value is replaced by zero, so no load waits behind a
// branch.  A point with no corner inside reads nothing (many such points would otherwise all
// read the same edge cell).
template <typename T>
__device__ inline Cplx<T> corner_sum(const Cplx<T>* __restrict__ v, const Corner8<T>& cn) {
  bool any = false;
#pragma unroll
  for (int k = 0; k < 8; ++k) any = any || cn.ok[k];
  if (!any) return {(T)0, (T)0};
  Cplx<T> a[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) a[k] = v[cn.off[k]];
  Cplx<T> acc = {(T)0, (T)0};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const T ax = cn.ok[k] ? a[k].x : (T)0, ay = cn.ok[k] ? a[k].y : (T)0;
    acc.x += ax * cn.w[k];
    acc.y += ay * cn.w[k];
  }
  return acc;
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
    for (int u = 0; u < kFwdGroup; ++u)
      if (t0 + u < ntrans) acc[u] = corner_sum(vol + (int64_t)(t0 + u) * vsize, cn);
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

// Point j's values of every transform (rows of length npts) into its row, times sc when scaled,
// 8 transforms' loads in flight at a time (a transform past the last rereads the last one's value
// and is not stored).
template <typename T>
__device__ inline void copy_row_values(const Cplx<T>* __restrict__ sig, Cplx<T>* __restrict__ row, int64_t j,
                                       int64_t npts, int ntrans, T sc = (T)1, bool scaled = false) {
  constexpr int kBatch = 8;
  for (int t0 = 0; t0 < ntrans; t0 += kBatch) {
    Cplx<T> v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) v[u] = sig[(int64_t)min(t0 + u, ntrans - 1) * npts + j];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      if (scaled) v[u] = {v[u].x * sc, v[u].y * sc};
      if (t0 + u < ntrans) row[2 + t0 + u] = v[u];
    }
  }
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
  copy_row_values(sig, row, j, npts, ntrans);
}

// Rows as k_slice_rows for a slice plan: the coordinates come from the point's sorted record
// (the plan keeps no grid pointer); the values are read in the caller's order (coalesced) and
// the row is written at the point's sorted position.  The values of the caller's points [0, n_a)
// come from sig_a (ntrans, n_a), the rest from sig_b (ntrans, npts - n_a); with a scale pointer
// every value is multiplied by *scale.
template <typename T>
__global__ void __launch_bounds__(kSliceBlock)
    k_slice_rows_plan(const Cplx<T>* __restrict__ sig_a, int64_t n_a, const Cplx<T>* __restrict__ sig_b,
                      const T* __restrict__ scale, const SortedPoint<T>* __restrict__ sp,
                      const int* __restrict__ spos, Cplx<T>* __restrict__ rows, int64_t npts, int ntrans, int rw) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npts) return;
  const int64_t s = spos[j];
  const SortedPoint<T> p = sp[s];
  Cplx<T>* row = rows + s * rw;
  row[0] = {p.x, p.y};
  row[1] = {p.z, (T)0};
  const bool in_a = j < n_a;
  copy_row_values(in_a ? sig_a : sig_b, row, in_a ? j : j - n_a, in_a ? n_a : npts - n_a, ntrans,
                  scale ? *scale : (T)1, scale != nullptr);
}

// k_slice_rows_plan through LDS, for rows of at most kRowsT complex words: a workgroup gathers
// the rows of 256 consecutive caller points (their values read coalesced, as above) into LDS and
// then writes them out cooperatively, consecutive lanes storing consecutive words of a row, so
// every row leaves as one contiguous stretch instead of as 8-byte pieces from 22 store
// instructions.
constexpr int kRowsT = 22;

template <typename T>
__global__ void __launch_bounds__(kSliceBlock)
    k_slice_rows_plan_lds(const Cplx<T>* __restrict__ sig_a, int64_t n_a, const Cplx<T>* __restrict__ sig_b,
                          const T* __restrict__ scale, const SortedPoint<T>* __restrict__ sp,
                          const int* __restrict__ spos, Cplx<T>* __restrict__ rows, int64_t npts, int ntrans,
                          int rw) {
  __shared__ Cplx<T> s_rows[kSliceBlock * kRowsT];
  __shared__ int s_pos[kSliceBlock];
  const int64_t j0 = (int64_t)blockIdx.x * kSliceBlock;
  const int64_t j = j0 + threadIdx.x;
  if (j < npts) {
    const int s = spos[j];
    const SortedPoint<T> p = sp[s];
    s_pos[threadIdx.x] = s;
    Cplx<T> row[kRowsT];
    row[0] = {p.x, p.y};
    row[1] = {p.z, (T)0};
    const bool in_a = j < n_a;
    const Cplx<T>* src = in_a ? sig_a : sig_b;
    const int64_t jj = in_a ? j : j - n_a, stride = in_a ? n_a : npts - n_a;
    const T sc = scale ? *scale : (T)1;
#pragma unroll
    for (int u = 0; u < kRowsT - 2; ++u) {
      Cplx<T> v = src[(int64_t)min(u, ntrans - 1) * stride + jj];
      if (scale) v = {v.x * sc, v.y * sc};
      row[2 + u] = v;
    }
#pragma unroll
    for (int u = 0; u < kRowsT; ++u)
      if (u < rw) s_rows[threadIdx.x * rw + u] = row[u];
  }
  __syncthreads();
  const int n = (int)min<int64_t>(kSliceBlock, npts - j0);
  const int ne = n * rw;
  for (int e = threadIdx.x; e < ne; e += kSliceBlock) {
    const int r = e / rw;
    rows[(int64_t)s_pos[r] * rw + (e - r * rw)] = s_rows[e];
  }
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
  const bool full = t0 + RT <= ntrans;
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
      if (full) {  // every load issued before the sums (a guarded load would wait behind its branch)
#pragma unroll
        for (int u = 0; u < RT; ++u) {
          const Cplx<T> v = row[2 + t0 + u];
          ax[u] += w * v.x;
          ay[u] += w * v.y;
        }
      } else {
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
  }
#pragma unroll
  for (int u = 0; u < RT; ++u)
    if (t0 + u < ntrans) vol[(int64_t)(t0 + u) * vsize + q] = {ax[u], ay[u]};
}

// Points per key (end - start, zero for absent keys), the input of the scan that gives off[].
__global__ void k_slice_counts(const int* __restrict__ start, const int* __restrict__ end, int* __restrict__ cnt,
                               int64_t nkeys) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nkeys) cnt[k] = end[k] - start[k];
  if (k == nkeys) cnt[k] = 0;
}

// Aggregated trilinear adjoint of a slice plan, tiled (FP32, ntrans <= 20).  A 256-thread
// workgroup owns a 16 x 4 x 4 (x, y, z) block of output cells.  The points whose floor cells touch
// the block, floor cells [x0 - 1, x0 + 15] x [y0 - 1, y0 + 3] x [z0 - 1, z0 + 3], are 25 runs of
// consecutive keys (one per (y, z) floor row), so their rows are 25 contiguous ranges of the sorted
// rows; taken in run order they form one list of ntot rows.  The list goes through LDS in chunks of
// 256 rows (each run's part of a chunk copied by the whole workgroup, coalesced), and every cell adds,
// chunk by chunk, the rows of its 8 floor cells that lie in the chunk, each times its trilinear
// weight (the products k_slice_adjoint_gather forms).  Each row is read from memory about 1.7 times
// (the halo of the block) instead of once per cell it reaches (8 times), and the sums run in a fixed
// order (chunks, then floor cells d in {0,1}^3, then sorted rows): deterministic, and equal to the
// per-cell gather to rounding.  A dense block (near the centre of the spectrum every slice passes)
// simply takes more chunks.
constexpr int kTileX = 16, kTileY = 4, kTileZ = 4;
constexpr int kTileRuns = (kTileY + 1) * (kTileZ + 1);
constexpr int kTileRowCap = kSliceBlock;  // rows per chunk: one per thread

template <int RT>
__global__ void __launch_bounds__(kSliceBlock)
    k_slice_adjoint_tile(const float2* __restrict__ rows, const int* __restrict__ off, float2* __restrict__ vol,
                         int64_t D, int64_t H, int64_t W, int ntrans, int rw, const int* __restrict__ border) {
  static_assert(kTileX * kTileY * kTileZ == kSliceBlock, "one cell per thread");
  __shared__ float2 s_rows[kTileRowCap * (2 + RT)];
  __shared__ int s_gstart[kTileRuns], s_len[kTileRuns], s_lbase[kTileRuns + 1];
  __shared__ int s_src[kTileRowCap];
  const int64_t nbx = W / kTileX, nby = H / kTileY;
  // blocks run nearest the volume centre first: the central blocks hold the most rows (slices
  // all pass through the centre), and the light outer ones then fill the end of the launch
  const int64_t blk = border[blockIdx.x];
  const int64_t x0 = (blk % nbx) * kTileX, y0 = ((blk / nbx) % nby) * kTileY, z0 = (blk / (nbx * nby)) * kTileZ;
  const int tx = threadIdx.x % kTileX, ty = (threadIdx.x / kTileX) % kTileY, tz = threadIdx.x / (kTileX * kTileY);
  auto key_of = [&](int64_t fx, int64_t fy, int64_t fz) { return ((fz + 1) * (H + 1) + (fy + 1)) * (W + 1) + (fx + 1); };
  if (threadIdx.x < kTileRuns) {
    const int64_t fy = y0 - 1 + threadIdx.x % (kTileY + 1), fz = z0 - 1 + threadIdx.x / (kTileY + 1);
    const int a = off[key_of(x0 - 1, fy, fz)], b = off[key_of(x0 + kTileX - 1, fy, fz) + 1];
    s_gstart[threadIdx.x] = a;
    s_len[threadIdx.x] = b - a;
  }
  __syncthreads();
  if (threadIdx.x <= kTileRuns) {  // exclusive prefix of the run lengths, one entry per thread
    int acc = 0;
    for (int i = 0; i < (int)threadIdx.x; ++i) acc += s_len[i];
    s_lbase[threadIdx.x] = acc;
  }
  __syncthreads();
  const int ntot = s_lbase[kTileRuns];
  const int64_t qx = x0 + tx, qy = y0 + ty, qz = z0 + tz;
  // this cell's 8 floor cells as ranges [lo, hi) of the run-ordered list
  int lo[8], hi[8];
#pragma unroll
  for (int nb = 0; nb < 8; ++nb) {
    const int dx = nb & 1, dy = (nb >> 1) & 1, dz = (nb >> 2) & 1;
    const int64_t key = key_of(qx - dx, qy - dy, qz - dz);
    const int run = (tz - dz + 1) * (kTileY + 1) + (ty - dy + 1);
    const int shift = s_lbase[run] - s_gstart[run];
    lo[nb] = off[key] + shift;
    hi[nb] = off[key + 1] + shift;
  }
  const float inv_rw = 1.f / (float)rw;
  float ax[RT], ay[RT];
#pragma unroll
  for (int u = 0; u < RT; ++u) ax[u] = ay[u] = 0.f;
  for (int c0 = 0; c0 < ntot; c0 += kTileRowCap) {
    const int n = min(kTileRowCap, ntot - c0);
    __syncthreads();  // the previous chunk is consumed
    if ((int)threadIdx.x < n) {
      // chunk row t is list entry g = c0 + t; run r holds entries [s_lbase[r], s_lbase[r + 1])
      const int gl = c0 + threadIdx.x;
      int r = 0;
#pragma unroll
      for (int step = 16; step >= 1; step >>= 1)
        if (r + step < kTileRuns && s_lbase[r + step] <= gl) r += step;
      s_src[threadIdx.x] = s_gstart[r] + (gl - s_lbase[r]);
    }
    __syncthreads();
    // the chunk as n * rw flat float2s, all loads issued before the first LDS write (a run's
    // rows are contiguous, so consecutive lanes read consecutive addresses)
    const int ne = n * rw;
    float2 v[2 + RT];
#pragma unroll
    for (int i = 0; i < 2 + RT; ++i) {
      const int e = threadIdx.x + i * kSliceBlock;
      if (e < ne) {
        const int row = (int)(((float)e + 0.5f) * inv_rw);  // exact: e + 0.5 is 0.5 / rw from a multiple
        v[i] = rows[(int64_t)s_src[row] * rw + (e - row * rw)];
      }
    }
#pragma unroll
    for (int i = 0; i < 2 + RT; ++i) {
      const int e = threadIdx.x + i * kSliceBlock;
      if (e < ne) s_rows[e] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
      const int dx = nb & 1, dy = (nb >> 1) & 1, dz = (nb >> 2) & 1;
      const int a = max(lo[nb], c0), b = min(hi[nb], c0 + n);
      for (int g = a; g < b; ++g) {
        const float2* row = s_rows + (g - c0) * rw;
        const float2 c0v = row[0], c1v = row[1];
        const float fx = floorf(c0v.x), fy = floorf(c0v.y), fz = floorf(c1v.x);
        const float wx = dx ? c0v.x - fx : (fx + 1.f) - c0v.x;
        const float wy = dy ? c0v.y - fy : (fy + 1.f) - c0v.y;
        const float wz = dz ? c1v.x - fz : (fz + 1.f) - c1v.x;
        const float w = wx * wy * wz;
        // all RT values read (beyond ntrans they are other LDS words, summed and never written)
#pragma unroll
        for (int u = 0; u < RT; ++u) {
          const float2 v = row[2 + u];
          ax[u] += w * v.x;
          ay[u] += w * v.y;
        }
      }
    }
  }
  const int64_t vsize = D * H * W;
  const int64_t q = (qz * H + qy) * W + qx;
#pragma unroll
  for (int u = 0; u < RT; ++u)
    if (u < ntrans) vol[(int64_t)u * vsize + q] = make_float2(ax[u], ay[u]);
}

// Sort workspace of the gather adjoint, per device, grown on demand and kept.
struct AdjWork {
  size_t cap_pts = 0, cap_keys = 0, cap_tmp = 0, cap_rows = 0;
  unsigned *keys = nullptr, *skeys = nullptr;
  int *sidx = nullptr, *spos = nullptr, *start = nullptr, *end = nullptr;
  int* off = nullptr;  // slice plans: first sorted position of every key (nkeys + 1 entries)
  int* cnt = nullptr;  // ... and the points per key it is scanned from
  size_t cap_off = 0;
  int* border = nullptr;  // tile blocks by distance from the volume centre, for border_shape
  int64_t border_shape[3] = {0, 0, 0};
  std::vector<int> border_host;  // the uploaded order (kept alive for the asynchronous copy)
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
// one thread per (sorted point, group of kFwdSortedTB transforms), the group the slow grid index.
// The caller's points [0, n_a) write out_a (ntrans, n_a), the rest out_b (ntrans, npts - n_a).
// Measured at L = 128, r = 20, vL = 256 (8.26 M points): 3.19 ms with one transform per thread,
// 3.27 with two, 3.55 with four (the point records are read once per group, but a thread's
// loads then span several 134 MB volumes).
constexpr int kFwdSortedTB = 1;  // transforms per thread

template <typename T, int MODE, int TB>
__global__ void __launch_bounds__(kSliceBlock)
    k_slice_forward_sorted(const Cplx<T>* __restrict__ vol, const SortedPoint<T>* __restrict__ sp,
                           Cplx<T>* __restrict__ out_a, int64_t n_a, Cplx<T>* __restrict__ out_b, int64_t D,
                           int64_t H, int64_t W, int64_t npts, int ntrans) {
  // (dealing consecutive sorted points to one XCD, xcd_contiguous, measured slower: 3.19 -> 3.41 ms)
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= npts) return;
  const int64_t t0 = (int64_t)blockIdx.y * TB;
  const int64_t vsize = D * H * W;
  const SortedPoint<T> p = sp[s];
  // one destination row per transform: the caller's point p.j in out_a or out_b
  const bool in_a = p.j < n_a;
  Cplx<T>* dst = in_a ? out_a + p.j : out_b + (p.j - n_a);
  const int64_t ostride = in_a ? n_a : npts - n_a;
  auto store = [&](int64_t t, Cplx<T> acc) {
    // one vector store (alternative member-wise stores were split into 4-byte pieces)
    if constexpr (sizeof(T) == 4)
      *reinterpret_cast<float2*>(dst + t * ostride) = make_float2((float)acc.x, (float)acc.y);
    else
      *reinterpret_cast<double2*>(dst + t * ostride) = make_double2((double)acc.x, (double)acc.y);
  };
  if (MODE == 0) {
    Corner8<T> cn;
    trilinear_corners(p.x, p.y, p.z, D, H, W, cn);
    if (t0 + TB <= ntrans) {
      // every transform's 8 loads in flight together (no per-transform branch)
      Cplx<T> acc[TB];
#pragma unroll
      for (int u = 0; u < TB; ++u) acc[u] = corner_sum(vol + (t0 + u) * vsize, cn);
#pragma unroll
      for (int u = 0; u < TB; ++u) store(t0 + u, acc[u]);
    } else {
      for (int64_t t = t0; t < ntrans; ++t) store(t, corner_sum(vol + t * vsize, cn));
    }
  } else {
    const int64_t x = (int64_t)rint(p.x), y = (int64_t)rint(p.y), z = (int64_t)rint(p.z);
    const bool in = x >= 0 && x < W && y >= 0 && y < H && z >= 0 && z < D;
    for (int64_t t = t0; t < t0 + TB && t < ntrans; ++t)
      store(t, in ? vol[t * vsize + (z * H + y) * W + x] : Cplx<T>{(T)0, (T)0});
  }
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
  // the key scan of the tiled adjoint shares the sort's work area: size it for both before either runs
  const bool want_off = sizeof(T) == 4 && mode == 0;
  size_t scan_bytes = 0;
  if (want_off) COV3D_HIP(exclusive_scan_ints(nullptr, scan_bytes, nullptr, nullptr, (int)(nkeys + 1), st));
  if (int rc = plan_grow(&w.tmp, w.cap_tmp, std::max(tmp_bytes, scan_bytes))) return rc;
  if (want_off && (size_t)(nkeys + 1) > w.cap_off) {
    hipFree(w.off), hipFree(w.cnt);
    w.off = w.cnt = nullptr, w.cap_off = 0;
    COV3D_HIP(hipMalloc(&w.off, (nkeys + 1) * sizeof(int)));
    COV3D_HIP(hipMalloc(&w.cnt, (nkeys + 1) * sizeof(int)));
    w.cap_off = nkeys + 1;
  }
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
  if (want_off) {
    // first sorted position of every key, for the tiled adjoint: an exclusive scan of the counts
    hipLaunchKernelGGL(k_slice_counts, dim3((unsigned)ceil_div(nkeys + 1, kSliceBlock)), dim3(kSliceBlock), 0, st,
                       w.start, w.end, w.cnt, nkeys);
    COV3D_LAUNCH_CHECK();
    size_t sb = w.cap_tmp;
    COV3D_HIP(exclusive_scan_ints(w.tmp, sb, w.cnt, w.off, (int)(nkeys + 1), st));
  }
  pl->npts = npts, pl->D = D, pl->H = H, pl->W = W, pl->nkeys = nkeys, pl->mode = mode;
  pl->ready = true;
  return 0;
}

template <typename T>
static int slice_plan_forward(SlicePlan* pl, const void* vol, void* out_a, int64_t n_a, void* out_b, int64_t ntrans,
                              hipStream_t st) {
  if (pl->npts == 0) return 0;
  dim3 g((unsigned)ceil_div(pl->npts, kSliceBlock), (unsigned)ceil_div(ntrans, kFwdSortedTB));
  if (pl->mode == 0)
    hipLaunchKernelGGL((k_slice_forward_sorted<T, 0, kFwdSortedTB>), g, dim3(kSliceBlock), 0, st,
                       (const Cplx<T>*)vol, (const SortedPoint<T>*)pl->sp, (Cplx<T>*)out_a, n_a, (Cplx<T>*)out_b,
                       pl->D, pl->H, pl->W, pl->npts, (int)ntrans);
  else
    hipLaunchKernelGGL((k_slice_forward_sorted<T, 1, kFwdSortedTB>), g, dim3(kSliceBlock), 0, st,
                       (const Cplx<T>*)vol, (const SortedPoint<T>*)pl->sp, (Cplx<T>*)out_a, n_a, (Cplx<T>*)out_b,
                       pl->D, pl->H, pl->W, pl->npts, (int)ntrans);
  COV3D_LAUNCH_CHECK();
  return 0;
}

template <typename T>
static int slice_plan_adjoint(SlicePlan* pl, const void* sig_a, int64_t n_a, const void* sig_b, const void* scale,
                              void* vol, int64_t ntrans, hipStream_t st) {
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
  if (rw <= kRowsT)
    hipLaunchKernelGGL((k_slice_rows_plan_lds<T>), dim3(nbp), dim3(kSliceBlock), 0, st, (const Cplx<T>*)sig_a,
                       n_a, (const Cplx<T>*)sig_b, (const T*)scale, (const SortedPoint<T>*)pl->sp, w.spos,
                       (Cplx<T>*)w.rows, npts, (int)ntrans, rw);
  else
    hipLaunchKernelGGL((k_slice_rows_plan<T>), dim3(nbp), dim3(kSliceBlock), 0, st, (const Cplx<T>*)sig_a, n_a,
                       (const Cplx<T>*)sig_b, (const T*)scale, (const SortedPoint<T>*)pl->sp, w.spos,
                       (Cplx<T>*)w.rows, npts, (int)ntrans, rw);
  COV3D_LAUNCH_CHECK();
  constexpr int RT = 20;
  dim3 g((unsigned)ceil_div(D * H * W, kSliceBlock), (unsigned)ceil_div(ntrans, RT));
  if (sizeof(T) == 4 && pl->mode == 0 && w.off && ntrans <= RT && W % kTileX == 0 && H % kTileY == 0 &&
      D % kTileZ == 0) {
    const int64_t nbx = W / kTileX, nby = H / kTileY, nbz = D / kTileZ;
    const unsigned nblk = (unsigned)(nbx * nby * nbz);
    if (w.border_shape[0] != D || w.border_shape[1] != H || w.border_shape[2] != W) {
      // block order by squared distance of the block centre from the volume centre (ties by index);
      // a previous order's upload must be finished before its host copy is rewritten
      if (w.border) COV3D_HIP(hipStreamSynchronize(st));
      std::vector<int64_t> key(nblk);
      w.border_host.resize(nblk);
      for (int64_t b = 0; b < (int64_t)nblk; ++b) {
        const int64_t bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
        const int64_t dx = 2 * (bx * kTileX) + kTileX - W, dy = 2 * (by * kTileY) + kTileY - H,
                      dz = 2 * (bz * kTileZ) + kTileZ - D;
        key[b] = dx * dx + dy * dy + dz * dz;
        w.border_host[b] = (int)b;
      }
      std::stable_sort(w.border_host.begin(), w.border_host.end(), [&](int a, int b) { return key[a] < key[b]; });
      hipFree(w.border);
      w.border = nullptr;
      w.border_shape[0] = w.border_shape[1] = w.border_shape[2] = 0;
      COV3D_HIP(hipMalloc(&w.border, (size_t)nblk * sizeof(int)));
      COV3D_HIP(hipMemcpyAsync(w.border, w.border_host.data(), (size_t)nblk * sizeof(int), hipMemcpyHostToDevice, st));
      w.border_shape[0] = D, w.border_shape[1] = H, w.border_shape[2] = W;
    }
    if (ntrans == 1)
      hipLaunchKernelGGL((k_slice_adjoint_tile<1>), dim3(nblk), dim3(kSliceBlock), 0, st, (const float2*)w.rows,
                         w.off, (float2*)vol, D, H, W, 1, rw, (const int*)w.border);
    else
      hipLaunchKernelGGL((k_slice_adjoint_tile<RT>), dim3(nblk), dim3(kSliceBlock), 0, st, (const float2*)w.rows,
                         w.off, (float2*)vol, D, H, W, (int)ntrans, rw, (const int*)w.border);
  } else if (pl->mode == 0)
    hipLaunchKernelGGL((k_slice_adjoint_gather<T, 0, RT>), g, dim3(kSliceBlock), 0, st, (const Cplx<T>*)w.rows,
                       w.start, w.end, (Cplx<T>*)vol, D, H, W, (int)ntrans, rw);
  else
    hipLaunchKernelGGL((k_slice_adjoint_gather<T, 1, RT>), g, dim3(kSliceBlock), 0, st, (const Cplx<T>*)w.rows,
                       w.start, w.end, (Cplx<T>*)vol, D, H, W, (int)ntrans, rw);
  COV3D_LAUNCH_CHECK();
  return 0;
}

// grid_sample coordinates of slice points from rotated frequencies (B, 3, npix) (zyx rows):
// grid[(b npix + k), c] = pts[b, 2 - c, k] * scale + shift, the per-point form of
// NufftPlanDiscretized.setpts' transpose, flip and affine map.
template <typename T>
__global__ void k_slice_grid_from_pts(const T* __restrict__ pts, T* __restrict__ grid, int64_t n, int64_t npix,
                                      T scale, T shift) {
  // a product and a sum, each rounded, as the two tensor operations round them (no fused multiply-add)
#pragma clang fp contract(off)
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // point: three coalesced loads
  if (j >= n * npix) return;
  const int64_t b = j / npix, k = j - b * npix;
  const T* pb = pts + b * 3 * npix + k;
  T out[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T a = pb[(2 - c) * npix] * scale;
    out[c] = shift != (T)0 ? a + shift : a;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) grid[3 * j + c] = out[c];
}

// Boundary pixels of the Fourier-domain ML on a slice plan: pixel k of image b is flagged when its
// trilinear stencil can reach cell 0 or cell vL of the padded volume on some axis (cell coordinate
// within one of 0 or vL), OR-ed with the flag of its mirror pixel -k, and kept where keep[k] is set
// (the half plane, or every pixel).  One workgroup per image writes the kept flagged pixels in
// ascending order to order[b, 0 .. count[b]) (at most cap of them; count[b] is the full number).
constexpr int kBndThreads = 256;

__device__ __forceinline__ bool near_nyquist(const float* g, float half_cells, float vL) {
  bool near = false;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float x = (g[c] + 1.f) * half_cells;
    near = near || (x > -1.f && x < 1.f) || (x > vL - 1.f && x < vL + 1.f);
  }
  return near;
}

// pass 1, one thread per pixel: flag[b, k] = kept and (k or its mirror reaches a Nyquist cell)
__global__ void k_slice_boundary_flags(const float* __restrict__ grid, const unsigned char* __restrict__ keep,
                                       int64_t B, int L, float vL, unsigned char* __restrict__ flag) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int np = L * L;
  if (e >= B * np) return;
  const int64_t b = e / np;
  const int k = (int)(e - b * np);
  unsigned char f = 0;
  if (keep[k]) {
    const int i = k / L, j = k - i * L;
    const int mi = (L % 2 == 0) ? (L - i) % L : L - 1 - i, mj = (L % 2 == 0) ? (L - j) % L : L - 1 - j;
    const float* gb = grid + b * (int64_t)np * 3;
    const float half_cells = 0.5f * (vL - 1.f);
    f = (near_nyquist(gb + (int64_t)k * 3, half_cells, vL) || near_nyquist(gb + (int64_t)(mi * L + mj) * 3, half_cells, vL))
            ? 1 : 0;
  }
  flag[e] = f;
}

// pass 2, one workgroup per image: the flagged pixels in ascending order (a block scan of the
// threads' counts over contiguous pixel runs read as bytes)
__global__ void __launch_bounds__(kBndThreads)
    k_slice_boundary(const unsigned char* __restrict__ flag, int L, int* __restrict__ order, int* __restrict__ count,
                     int cap) {
  __shared__ int s_scan[kBndThreads];
  const int64_t b = blockIdx.x;
  const int np = L * L;
  const unsigned char* fb = flag + b * (int64_t)np;
  const int per = (np + kBndThreads - 1) / kBndThreads;
  const int k0 = threadIdx.x * per, k1 = min(np, k0 + per);
  int n = 0;
  for (int k = k0; k < k1; ++k) n += fb[k];
  s_scan[threadIdx.x] = n;
  __syncthreads();
  // inclusive scan (Hillis-Steele) over the threads' counts
  for (int off = 1; off < kBndThreads; off <<= 1) {
    const int v = threadIdx.x >= off ? s_scan[threadIdx.x - off] : 0;
    __syncthreads();
    s_scan[threadIdx.x] += v;
    __syncthreads();
  }
  int pos = s_scan[threadIdx.x] - n;
  if (n > 0) {
    for (int k = k0; k < k1; ++k)
      if (fb[k]) {
        if (pos < cap) order[b * cap + pos] = k;
        ++pos;
      }
  }
  if (threadIdx.x == kBndThreads - 1) count[b] = s_scan[kBndThreads - 1];
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

extern "C" int cov3d_slice_plan_forward_split(void* plan, const void* vol, void* out_a, int64_t n_a, void* out_b,
                                              int64_t ntrans, void* stream) {
  COV3D_CHECK_ARG(plan, "NULL plan");
  auto* pl = static_cast<cov3d::SlicePlan*>(plan);
  COV3D_CHECK_ARG(pl->ready, "slice plan: setpts first");
  COV3D_CHECK_ARG(ntrans >= 1 && ntrans < 65536, "bad ntrans");
  COV3D_CHECK_ARG(n_a >= 0 && n_a <= pl->npts, "n_a must lie in [0, npts]");
  if (pl->npts == 0) return 0;
  COV3D_CHECK_ARG(vol && (out_a || n_a == 0) && (out_b || n_a == pl->npts), "NULL pointer");
  hipStream_t st = cov3d::as_stream(stream);
  if (pl->dtype == 1) return cov3d::slice_plan_forward<double>(pl, vol, out_a, n_a, out_b, ntrans, st);
  return cov3d::slice_plan_forward<float>(pl, vol, out_a, n_a, out_b, ntrans, st);
}

extern "C" int cov3d_slice_plan_forward(void* plan, const void* vol, void* out, int64_t ntrans, void* stream) {
  COV3D_CHECK_ARG(plan, "NULL plan");
  return cov3d_slice_plan_forward_split(plan, vol, out, static_cast<cov3d::SlicePlan*>(plan)->npts, nullptr, ntrans,
                                        stream);
}

extern "C" int cov3d_slice_plan_adjoint_split(void* plan, const void* sig_a, int64_t n_a, const void* sig_b,
                                              const void* scale, void* vol_out, int64_t ntrans, void* stream) {
  COV3D_CHECK_ARG(plan, "NULL plan");
  auto* pl = static_cast<cov3d::SlicePlan*>(plan);
  COV3D_CHECK_ARG(pl->ready, "slice plan: setpts first");
  COV3D_CHECK_ARG(ntrans >= 1 && ntrans < 65536, "bad ntrans");
  COV3D_CHECK_ARG(n_a >= 0 && n_a <= pl->npts, "n_a must lie in [0, npts]");
  COV3D_CHECK_ARG(vol_out && (sig_a || n_a == 0) && (sig_b || n_a == pl->npts), "NULL pointer");
  hipStream_t st = cov3d::as_stream(stream);
  if (pl->dtype == 1) return cov3d::slice_plan_adjoint<double>(pl, sig_a, n_a, sig_b, scale, vol_out, ntrans, st);
  return cov3d::slice_plan_adjoint<float>(pl, sig_a, n_a, sig_b, scale, vol_out, ntrans, st);
}

extern "C" int cov3d_slice_plan_adjoint(void* plan, const void* sig, void* vol_out, int64_t ntrans, void* stream) {
  COV3D_CHECK_ARG(plan, "NULL plan");
  return cov3d_slice_plan_adjoint_split(plan, sig, static_cast<cov3d::SlicePlan*>(plan)->npts, nullptr, nullptr,
                                        vol_out, ntrans, stream);
}

extern "C" int cov3d_slice_plan_destroy(void* plan) {
  if (!plan) return 0;
  auto* pl = static_cast<cov3d::SlicePlan*>(plan);
  cov3d::AdjWork& w = pl->w;
  hipFree(w.keys), hipFree(w.skeys), hipFree(w.sidx), hipFree(w.spos), hipFree(w.start), hipFree(w.end);
  hipFree(w.tmp), hipFree(w.rows), hipFree(w.off), hipFree(w.cnt), hipFree(w.border), hipFree(pl->sp);
  delete pl;
  return 0;
}

extern "C" int cov3d_slice_grid_from_pts(const void* pts, void* grid, int64_t n, int64_t npix, double scale,
                                         double shift, int dtype, void* stream) {
  COV3D_CHECK_ARG(n >= 0 && npix >= 0, "negative size");
  COV3D_CHECK_ARG(dtype == 0 || dtype == 1, "dtype must be 0 (float32) or 1 (float64)");
  if (n == 0 || npix == 0) return 0;
  COV3D_CHECK_ARG(pts && grid, "NULL pointer");
  hipStream_t st = cov3d::as_stream(stream);
  const unsigned nb = (unsigned)cov3d::ceil_div(n * npix, cov3d::kSliceBlock);
  if (dtype == 1)
    hipLaunchKernelGGL((cov3d::k_slice_grid_from_pts<double>), dim3(nb), dim3(cov3d::kSliceBlock), 0, st,
                       (const double*)pts, (double*)grid, n, npix, scale, shift);
  else
    hipLaunchKernelGGL((cov3d::k_slice_grid_from_pts<float>), dim3(nb), dim3(cov3d::kSliceBlock), 0, st,
                       (const float*)pts, (float*)grid, n, npix, (float)scale, (float)shift);
  COV3D_LAUNCH_CHECK();
  return 0;
}

extern "C" int cov3d_slice_boundary_pixels(const float* grid, const unsigned char* keep, int64_t B, int64_t L,
                                           int64_t vL, unsigned char* flags, int* order, int* count, int64_t cap,
                                           void* stream) {
  COV3D_CHECK_ARG(B >= 0 && L >= 1 && vL >= 2 && cap >= 0, "bad sizes");
  COV3D_CHECK_ARG(B < 2147483647 && L * L < 2147483647 && cap < 2147483647, "sizes too large");
  if (B == 0) return 0;
  COV3D_CHECK_ARG(grid && keep && flags && count && (order || cap == 0), "NULL pointer");
  hipStream_t st = cov3d::as_stream(stream);
  const int64_t npx = B * L * L;
  hipLaunchKernelGGL(cov3d::k_slice_boundary_flags, dim3((unsigned)cov3d::ceil_div(npx, cov3d::kSliceBlock)),
                     dim3(cov3d::kSliceBlock), 0, st, grid, keep, B, (int)L, (float)vL, flags);
  COV3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(cov3d::k_slice_boundary, dim3((unsigned)B), dim3(cov3d::kBndThreads), 0, st,
                     (const unsigned char*)flags, (int)L, order, count, (int)cap);
  COV3D_LAUNCH_CHECK();
  return 0;
}
