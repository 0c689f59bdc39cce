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
#include "common.h"

namespace cov3d {

constexpr int kSliceBlock = 256;

template <typename T>
struct Corner8 {
  int64_t off[8];
  T w[8];
  bool ok[8];
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

// One thread per (point, trans); trans is the slow grid dimension so the
// point coordinates are shared through L1/L2 across trans.
template <typename T, int MODE>
__global__ void __launch_bounds__(kSliceBlock)
    k_slice_forward(const Cplx<T>* __restrict__ vol, const T* __restrict__ grid,
                    Cplx<T>* __restrict__ out, int64_t D, int64_t H, int64_t W, int64_t npts) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t t = blockIdx.y;
  if (j >= npts) return;
  const Cplx<T>* v = vol + t * D * H * W;
  T ix = unnorm(grid[3 * j], W), iy = unnorm(grid[3 * j + 1], H), iz = unnorm(grid[3 * j + 2], D);
  Cplx<T> acc = {(T)0, (T)0};
  if (MODE == 0) {
    Corner8<T> cn;
    trilinear_corners(ix, iy, iz, D, H, W, cn);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (cn.ok[k]) {
        Cplx<T> a = v[cn.off[k]];
        acc.x += a.x * cn.w[k];
        acc.y += a.y * cn.w[k];
      }
    }
  } else {
    int64_t x = (int64_t)rint(ix), y = (int64_t)rint(iy), z = (int64_t)rint(iz);
    if (x >= 0 && x < W && y >= 0 && y < H && z >= 0 && z < D) acc = v[(z * H + y) * W + x];
  }
  out[t * npts + j] = acc;
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

template <typename T>
static int slice_fwd(const void* vol, const void* grid, void* out, int64_t ntrans, int64_t D,
                     int64_t H, int64_t W, int64_t npts, int mode, hipStream_t st) {
  dim3 g((unsigned)ceil_div(npts, kSliceBlock), (unsigned)ntrans);
  if (mode == 0)
    hipLaunchKernelGGL((k_slice_forward<T, 0>), g, dim3(kSliceBlock), 0, st, (const Cplx<T>*)vol,
                       (const T*)grid, (Cplx<T>*)out, D, H, W, npts);
  else
    hipLaunchKernelGGL((k_slice_forward<T, 1>), g, dim3(kSliceBlock), 0, st, (const Cplx<T>*)vol,
                       (const T*)grid, (Cplx<T>*)out, D, H, W, npts);
  COV3D_LAUNCH_CHECK();
  return 0;
}

template <typename T>
static int slice_adj(const void* sig, const void* grid, void* vol, int64_t ntrans, int64_t D,
                     int64_t H, int64_t W, int64_t npts, int mode, int aggregated, int64_t ppv,
                     hipStream_t st) {
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
