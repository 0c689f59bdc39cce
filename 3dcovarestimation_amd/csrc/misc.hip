// Data-feed helpers (rotated slice frequencies, CTF), FP64 complex dot
// products of the mean-volume CG, and the small MLP of the learned CG
// preconditioner.  The CG's Toeplitz operator lives in toeplitz.hip.
#include <math.h>

#include <map>
#include <mutex>
#include <utility>

#include "common.h"

namespace cov3d {

constexpr int kMiscBlock = 256;

// ----------------------------------------------------------------------------
// pts_rot: p = wrap(flip_zyx(R . pi*[gx, gy, 0])), grid_2d(L, "yx") values
// g_k = ceil(k - L/2) / (L/2) (x along columns, y along rows).
// Follows PoseModule._init_grid/forward (cov3d/poses.py:197-269).
// ----------------------------------------------------------------------------
__global__ void k_pts_rot(const float* __restrict__ rot, float* __restrict__ pts, int64_t B,
                          int L) {
  const int64_t d = (int64_t)L * L;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * d) return;
  const int64_t i = idx / d;
  const int64_t k = idx - i * d;
  const int row = (int)(k / L), col = (int)(k % L);
  const float half = 0.5f * (float)L;
  const float gx = ceilf((float)col - half) / half;
  const float gy = ceilf((float)row - half) / half;
  const float px = (float)M_PI * gx, py = (float)M_PI * gy;
  const float* R = rot + 9 * i;
  float v[3];
  for (int a = 0; a < 3; ++a) v[a] = R[3 * a] * px + R[3 * a + 1] * py;
  const float twopi = 2.f * (float)M_PI;
  for (int a = 0; a < 3; ++a) {
    float p = v[2 - a] + (float)M_PI;
    p = p - twopi * floorf(p / twopi);  // torch.remainder semantics
    pts[(i * 3 + a) * d + k] = p - (float)M_PI;
  }
}

// Points (n, 3, P) -> (3, n P) wrapped to [-pi, pi): remainder(x + pi, 2 pi) - pi, the
// torch.remainder formula x - m floor(x / m), in the points' own precision.
template <typename T>
__global__ void k_points_wrap(const T* __restrict__ in, T* __restrict__ out, int64_t n, int64_t P) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * P) return;
  const int64_t i = idx / P, k = idx - i * P;
  const T pi = (T)M_PI, twopi = (T)(2.0 * M_PI);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const T p = in[(i * 3 + a) * P + k] + pi;
    out[a * n * P + idx] = (p - twopi * floor(p / twopi)) - pi;
  }
}

// ----------------------------------------------------------------------------
// CTF (the formula of aspire's CTFFilter._evaluate on evaluate_grid's grid):
//   omega = pi*[gx, gy]; f = omega / (2 pi pixel); angle = atan2(fy, fx) - ang;
//   defocus = mean + diff*cos(2 angle) (0 at f = 0);
//   gamma = -pi*lambda*defocus*|f|^2 + 0.5*pi*Cs*1e7*lambda^3*|f|^4;
//   h = sqrt(1 - a^2) sin(gamma) - a cos(gamma).
// ----------------------------------------------------------------------------
__global__ void k_ctf(const float* __restrict__ prm, float* __restrict__ out, int64_t B, int L) {
  const int64_t d = (int64_t)L * L;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * d) return;
  const int64_t i = idx / d;
  const int64_t k = idx - i * d;
  const int row = (int)(k / L), col = (int)(k % L);
  const float* p = prm + 7 * i;
  const double du = p[0], dv = p[1], ang = p[2], kv = p[3], cs = p[4], alpha = p[5], px = p[6];
  const double half = 0.5 * L;
  const double gx = ceil((double)col - half) / half, gy = ceil((double)row - half) / half;
  const double fx = M_PI * gx / (2 * M_PI * px), fy = M_PI * gy / (2 * M_PI * px);
  const double lam = 12.2643247 / sqrt(kv * 1e3 + 0.978466 * kv * kv);
  const double eps = 2.220446049250313e-16;
  double defocus = 0;
  if (fabs(fx) > eps || fabs(fy) > eps) {
    double a = atan2(fy, fx) - ang;
    defocus = 0.5 * (du + dv) + 0.5 * (du - dv) * cos(2 * a);
  }
  const double r2 = fx * fx + fy * fy;
  const double gamma = -M_PI * lam * defocus * r2 + 0.5 * M_PI * (cs * 1e7) * lam * lam * lam * r2 * r2;
  out[idx] = (float)(sqrt(1 - alpha * alpha) * sin(gamma) - alpha * cos(gamma));
}

// ----------------------------------------------------------------------------
// FP64-accumulated complex dot product sum conj(x) y.
// ----------------------------------------------------------------------------
__global__ void k_zero2(double* out) {
  out[0] = 0;
  out[1] = 0;
}

__global__ void __launch_bounds__(kMiscBlock)
    k_vdot(const float2* __restrict__ x, const float2* __restrict__ y, int64_t n,
           double* __restrict__ out) {
  __shared__ double sr[kMiscBlock], si[kMiscBlock];
  // kDotBatch elements per thread per step with all their loads issued first (one element per
  // step leaves one load pair in flight per thread: 256 blocks then wait on latency), the rest
  // one at a time
  constexpr int kDotBatch = 8;
  double ar = 0, ai = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; k + (kDotBatch - 1) * stride < n; k += kDotBatch * stride) {
    float2 a[kDotBatch], b[kDotBatch];
#pragma unroll
    for (int u = 0; u < kDotBatch; ++u) {
      a[u] = x[k + u * stride];
      b[u] = y[k + u * stride];
    }
#pragma unroll
    for (int u = 0; u < kDotBatch; ++u) {
      ar += (double)a[u].x * b[u].x + (double)a[u].y * b[u].y;
      ai += (double)a[u].x * b[u].y - (double)a[u].y * b[u].x;
    }
  }
  for (; k < n; k += stride) {
    const float2 a = x[k], b = y[k];
    ar += (double)a.x * b.x + (double)a.y * b.y;
    ai += (double)a.x * b.y - (double)a.y * b.x;
  }
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int s = kMiscBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sr[threadIdx.x] += sr[threadIdx.x + s];
      si[threadIdx.x] += si[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(out, sr[0]);
    atomicAdd(out + 1, si[0]);
  }
}

// One FP64 atomic pair per block lands on the same two addresses, so the reductions use at most
// one block per CU: 2048 blocks (4096 contended atomics) took 54 us for 2 M elements, the loads
// alone take a few.
__host__ __device__ inline int64_t dot_blocks(int64_t n) {
  const int64_t nb = (n + kMiscBlock - 1) / kMiscBlock;
  return nb < 256 ? (nb < 1 ? 1 : nb) : 256;
}

// Block sum of (a, b) added to out[0], out[1].
__device__ __forceinline__ void block_add2(double a, double b, double* out) {
  __shared__ double sa[kMiscBlock], sb[kMiscBlock];
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int s = kMiscBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sa[threadIdx.x] += sa[threadIdx.x + s];
      sb[threadIdx.x] += sb[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(out, sa[0]);
    atomicAdd(out + 1, sb[0]);
  }
}

// CG step x += alpha p, r -= alpha Ap with alpha = rz / pAp (complex, from the FP64 device
// scalars: rz real, pAp = vdot(p, Ap)), rounded to complex64 as the reference's complex64
// alpha; rr += |r|^2 (FP64) of the updated residual.
__global__ void __launch_bounds__(kMiscBlock)
    k_cg_update(float2* __restrict__ x, float2* __restrict__ r, const float2* __restrict__ p,
                const float2* __restrict__ ap, int64_t n, const double* __restrict__ rz,
                const double* __restrict__ pap, double* __restrict__ rr) {
  const double den = pap[0] * pap[0] + pap[1] * pap[1];
  const float ar = (float)(rz[0] * pap[0] / den), ai = (float)(-rz[0] * pap[1] / den);
  double acc = 0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const float2 pk = p[k], qk = ap[k];
    float2 xk = x[k], rk = r[k];
    const float2 tp = make_float2(ar * pk.x - ai * pk.y, ar * pk.y + ai * pk.x);
    const float2 tq = make_float2(ar * qk.x - ai * qk.y, ar * qk.y + ai * qk.x);
    xk.x += tp.x;
    xk.y += tp.y;
    rk.x -= tq.x;
    rk.y -= tq.y;
    x[k] = xk;
    r[k] = rk;
    acc += (double)rk.x * rk.x + (double)rk.y * rk.y;
  }
  block_add2(acc, 0.0, rr);
}

// CG direction p = beta p + z with beta = rz_new / rz_old (real parts of FP64 device scalars),
// rounded to complex64 as the reference's complex64 beta.
__global__ void __launch_bounds__(kMiscBlock)
    k_cg_direction(float2* __restrict__ p, const float2* __restrict__ z, int64_t n, const double* __restrict__ rz_new,
                   const double* __restrict__ rz_old) {
  const float beta = (float)(rz_new[0] / rz_old[0]);
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float2 pk = p[k], zk = z[k];
  p[k] = make_float2(pk.x * beta + zk.x, pk.y * beta + zk.y);
}

// ----------------------------------------------------------------------------
// MLP (n_in -> 64 -> 64 -> 64 -> 1, GELU tanh form), one thread per row.
// ----------------------------------------------------------------------------
constexpr int kHid = 64;
constexpr int kMaxIn = 8;

__device__ __forceinline__ float gelu(float a) {
  const float c = 0.7978845608028654f;  // sqrt(2/pi)
  float u = c * (a + 0.044715f * a * a * a);
  return 0.5f * a * (1.f + tanhf(u));
}
__device__ __forceinline__ float gelu_grad(float a) {
  const float c = 0.7978845608028654f;
  float u = c * (a + 0.044715f * a * a * a);
  float t = tanhf(u);
  float du = c * (1.f + 3.f * 0.044715f * a * a);
  return 0.5f * (1.f + t) + 0.5f * a * (1.f - t * t) * du;
}

__host__ __device__ inline int64_t mlp_nparams(int n_in) {
  return (int64_t)kHid * n_in + kHid + 2 * (kHid * kHid + kHid) + kHid + 1;
}

struct MlpOffsets {
  int w1, b1, w2, b2, w3, b3, w4, b4;
};
__host__ __device__ inline MlpOffsets mlp_offsets(int n_in) {
  MlpOffsets o;
  o.w1 = 0;
  o.b1 = o.w1 + kHid * n_in;
  o.w2 = o.b1 + kHid;
  o.b2 = o.w2 + kHid * kHid;
  o.w3 = o.b2 + kHid;
  o.b3 = o.w3 + kHid * kHid;
  o.w4 = o.b3 + kHid;
  o.b4 = o.w4 + kHid;
  return o;
}

// Forward of one row; a1..a3 receive the pre-activations when non-null.
__device__ float mlp_row(const float* P, int n_in, const float* x, float* a1, float* a2,
                         float* a3) {
  MlpOffsets o = mlp_offsets(n_in);
  float h1[kHid], h2[kHid];
  for (int j = 0; j < kHid; ++j) {
    float s = P[o.b1 + j];
    for (int q = 0; q < n_in; ++q) s += P[o.w1 + j * n_in + q] * x[q];
    if (a1) a1[j] = s;
    h1[j] = gelu(s);
  }
  for (int j = 0; j < kHid; ++j) {
    float s = P[o.b2 + j];
    for (int q = 0; q < kHid; ++q) s += P[o.w2 + j * kHid + q] * h1[q];
    if (a2) a2[j] = s;
    h2[j] = gelu(s);
  }
  float out = P[o.b4];
  for (int j = 0; j < kHid; ++j) {
    float s = P[o.b3 + j];
    for (int q = 0; q < kHid; ++q) s += P[o.w3 + j * kHid + q] * h2[q];
    if (a3) a3[j] = s;
    out += P[o.w4 + j] * gelu(s);
  }
  return out;
}

__global__ void __launch_bounds__(kMiscBlock)
    k_mlp_forward(const float* __restrict__ params, const float* __restrict__ x,
                  float* __restrict__ y, int64_t n, int n_in) {
  extern __shared__ float sP[];
  const int64_t np = mlp_nparams(n_in);
  for (int64_t e = threadIdx.x; e < np; e += blockDim.x) sP[e] = params[e];
  __syncthreads();
  int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  float xi[kMaxIn];
  for (int q = 0; q < n_in; ++q) xi[q] = x[row * n_in + q];
  y[row] = mlp_row(sP, n_in, xi, nullptr, nullptr, nullptr);
}

// Per-row backward; parameter gradients are summed in LDS then added to
// grad_params with one atomic per parameter per block.
__global__ void __launch_bounds__(kMiscBlock)
    k_mlp_backward(const float* __restrict__ params, const float* __restrict__ x,
                   const float* __restrict__ gy, float* __restrict__ gparams, int64_t n,
                   int n_in) {
  extern __shared__ float smem[];
  const int64_t np = mlp_nparams(n_in);
  float* sP = smem;
  float* sG = smem + np;
  for (int64_t e = threadIdx.x; e < np; e += blockDim.x) {
    sP[e] = params[e];
    sG[e] = 0.f;
  }
  __syncthreads();
  int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row < n) {
    MlpOffsets o = mlp_offsets(n_in);
    float xi[kMaxIn];
    for (int q = 0; q < n_in; ++q) xi[q] = x[row * n_in + q];
    float a1[kHid], a2[kHid], a3[kHid];
    (void)mlp_row(sP, n_in, xi, a1, a2, a3);
    const float g = gy[row];
    atomicAdd(&sG[o.b4], g);
    float d3[kHid];
    for (int j = 0; j < kHid; ++j) {
      atomicAdd(&sG[o.w4 + j], g * gelu(a3[j]));
      d3[j] = g * sP[o.w4 + j] * gelu_grad(a3[j]);
    }
    // layer 3: s3 = W3 h2 + b3
    float d2[kHid];
    for (int q = 0; q < kHid; ++q) d2[q] = 0.f;
    for (int j = 0; j < kHid; ++j) {
      atomicAdd(&sG[o.b3 + j], d3[j]);
      for (int q = 0; q < kHid; ++q) {
        atomicAdd(&sG[o.w3 + j * kHid + q], d3[j] * gelu(a2[q]));
        d2[q] += sP[o.w3 + j * kHid + q] * d3[j];
      }
    }
    for (int q = 0; q < kHid; ++q) d2[q] *= gelu_grad(a2[q]);
    float d1[kHid];
    for (int q = 0; q < kHid; ++q) d1[q] = 0.f;
    for (int j = 0; j < kHid; ++j) {
      atomicAdd(&sG[o.b2 + j], d2[j]);
      for (int q = 0; q < kHid; ++q) {
        atomicAdd(&sG[o.w2 + j * kHid + q], d2[j] * gelu(a1[q]));
        d1[q] += sP[o.w2 + j * kHid + q] * d2[j];
      }
    }
    for (int j = 0; j < kHid; ++j) {
      float dj = d1[j] * gelu_grad(a1[j]);
      atomicAdd(&sG[o.b1 + j], dj);
      for (int q = 0; q < n_in; ++q) atomicAdd(&sG[o.w1 + j * n_in + q], dj * xi[q]);
    }
  }
  __syncthreads();
  for (int64_t e = threadIdx.x; e < np; e += blockDim.x)
    if (sG[e] != 0.f) atomicAdd(&gparams[e], sG[e]);
}

}  // namespace cov3d

extern "C" int cov3d_pts_rot(const void* rot, void* pts, int64_t B, int64_t L, void* stream) {
  COV3D_CHECK_ARG(B >= 0 && L >= 1, "bad shape");
  if (B == 0) return 0;
  COV3D_CHECK_ARG(rot && pts, "NULL pointer");
  int64_t n = B * L * L;
  hipLaunchKernelGGL(cov3d::k_pts_rot, dim3((unsigned)cov3d::ceil_div(n, cov3d::kMiscBlock)),
                     dim3(cov3d::kMiscBlock), 0, cov3d::as_stream(stream), (const float*)rot,
                     (float*)pts, B, (int)L);
  COV3D_LAUNCH_CHECK();
  return 0;
}

extern "C" int cov3d_points_wrap(const void* in, void* out, int64_t n, int64_t P, int dtype, void* stream) {
  COV3D_CHECK_ARG(n >= 0 && P >= 0, "bad shape");
  COV3D_CHECK_ARG(dtype == 0 || dtype == 1, "dtype must be 0 (float32) or 1 (float64)");
  if (n * P == 0) return 0;
  COV3D_CHECK_ARG(in && out && in != out, "NULL or aliased pointer");
  const dim3 grid((unsigned)cov3d::ceil_div(n * P, cov3d::kMiscBlock));
  if (dtype == 0)
    hipLaunchKernelGGL(cov3d::k_points_wrap<float>, grid, dim3(cov3d::kMiscBlock), 0, cov3d::as_stream(stream),
                       (const float*)in, (float*)out, n, P);
  else
    hipLaunchKernelGGL(cov3d::k_points_wrap<double>, grid, dim3(cov3d::kMiscBlock), 0, cov3d::as_stream(stream),
                       (const double*)in, (double*)out, n, P);
  COV3D_LAUNCH_CHECK();
  return 0;
}

extern "C" int cov3d_ctf_eval(const void* params, void* out, int64_t B, int64_t L, void* stream) {
  COV3D_CHECK_ARG(B >= 0 && L >= 1, "bad shape");
  if (B == 0) return 0;
  COV3D_CHECK_ARG(params && out, "NULL pointer");
  int64_t n = B * L * L;
  hipLaunchKernelGGL(cov3d::k_ctf, dim3((unsigned)cov3d::ceil_div(n, cov3d::kMiscBlock)),
                     dim3(cov3d::kMiscBlock), 0, cov3d::as_stream(stream), (const float*)params,
                     (float*)out, B, (int)L);
  COV3D_LAUNCH_CHECK();
  return 0;
}

extern "C" int cov3d_vdot_c64_f64(const void* x, const void* y, int64_t n, void* out,
                                  void* stream) {
  COV3D_CHECK_ARG(n >= 0, "bad length");
  COV3D_CHECK_ARG(out, "NULL output");
  hipStream_t st = cov3d::as_stream(stream);
  hipLaunchKernelGGL(cov3d::k_zero2, dim3(1), dim3(1), 0, st, (double*)out);
  COV3D_LAUNCH_CHECK();
  if (n == 0) return 0;
  COV3D_CHECK_ARG(x && y, "NULL input");
  hipLaunchKernelGGL(cov3d::k_vdot, dim3((unsigned)cov3d::dot_blocks(n)), dim3(cov3d::kMiscBlock), 0, st,
                     (const float2*)x, (const float2*)y, n, (double*)out);
  COV3D_LAUNCH_CHECK();
  return 0;
}

extern "C" int cov3d_mlp_forward(const void* params, const void* x, void* y, int64_t n,
                                 int64_t n_in, void* stream) {
  COV3D_CHECK_ARG(n_in >= 1 && n_in <= cov3d::kMaxIn, "n_in must be in [1, 8]");
  COV3D_CHECK_ARG(n >= 0, "bad row count");
  if (n == 0) return 0;
  COV3D_CHECK_ARG(params && x && y, "NULL pointer");
  size_t smem = (size_t)cov3d::mlp_nparams((int)n_in) * sizeof(float);
  hipLaunchKernelGGL(cov3d::k_mlp_forward, dim3((unsigned)cov3d::ceil_div(n, cov3d::kMiscBlock)),
                     dim3(cov3d::kMiscBlock), smem, cov3d::as_stream(stream),
                     (const float*)params, (const float*)x, (float*)y, n, (int)n_in);
  COV3D_LAUNCH_CHECK();
  return 0;
}

extern "C" int cov3d_mlp_backward(const void* params, const void* x, const void* gy,
                                  void* grad_params, int64_t n, int64_t n_in, void* stream) {
  COV3D_CHECK_ARG(n_in >= 1 && n_in <= cov3d::kMaxIn, "n_in must be in [1, 8]");
  COV3D_CHECK_ARG(n >= 0, "bad row count");
  if (n == 0) return 0;
  COV3D_CHECK_ARG(params && x && gy && grad_params, "NULL pointer");
  size_t smem = (size_t)2 * cov3d::mlp_nparams((int)n_in) * sizeof(float);
  static bool attr = false;
  if (!attr) {
    COV3D_HIP(hipFuncSetAttribute((const void*)cov3d::k_mlp_backward,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr = true;
  }
  hipLaunchKernelGGL(cov3d::k_mlp_backward, dim3((unsigned)cov3d::ceil_div(n, cov3d::kMiscBlock)),
                     dim3(cov3d::kMiscBlock), smem, cov3d::as_stream(stream),
                     (const float*)params, (const float*)x, (const float*)gy, (float*)grad_params,
                     n, (int)n_in);
  COV3D_LAUNCH_CHECK();
  return 0;
}

extern "C" int cov3d_cg_update(void* x, void* r, const void* p, const void* ap, int64_t n, const void* rz,
                               const void* pap, void* rr, void* stream) {
  COV3D_CHECK_ARG(n >= 1, "bad length");
  COV3D_CHECK_ARG(x && r && p && ap && rz && pap && rr, "NULL pointer");
  hipStream_t st = cov3d::as_stream(stream);
  hipLaunchKernelGGL(cov3d::k_zero2, dim3(1), dim3(1), 0, st, (double*)rr);
  hipLaunchKernelGGL(cov3d::k_cg_update, dim3((unsigned)cov3d::dot_blocks(n)), dim3(cov3d::kMiscBlock), 0, st,
                     (float2*)x, (float2*)r, (const float2*)p, (const float2*)ap, n, (const double*)rz,
                     (const double*)pap, (double*)rr);
  COV3D_LAUNCH_CHECK();
  return 0;
}

extern "C" int cov3d_cg_direction(void* p, const void* z, int64_t n, const void* rz_new, const void* rz_old,
                                  void* stream) {
  COV3D_CHECK_ARG(n >= 1, "bad length");
  COV3D_CHECK_ARG(p && z && rz_new && rz_old, "NULL pointer");
  hipLaunchKernelGGL(cov3d::k_cg_direction, dim3((unsigned)cov3d::ceil_div(n, cov3d::kMiscBlock)),
                     dim3(cov3d::kMiscBlock), 0, cov3d::as_stream(stream), (float2*)p, (const float2*)z, n,
                     (const double*)rz_new, (const double*)rz_old);
  COV3D_LAUNCH_CHECK();
  return 0;
}
