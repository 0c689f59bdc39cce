// Pruned complex Toeplitz operator of the mean-volume CG (FP32, cubic grids of NF = 64, 128, 256
// or 512 cells per axis, L <= NF):
//   y = crop_L(ifft3c(lhs * fft3c(pad_NF(x)))),   x, y: (L, L, L) complex,  lhs: (NF, NF, NF) complex.
// In uncentred indices (per axis) the padding puts x[i] at (i - L/2) mod NF, the multiply takes
// lhs[(k + NF/2) mod NF] at frequency k, and the crop reads y[i] from (i - L/2) mod NF / NF^3.
//
// Only L of NF inputs are non-zero on each axis of the forward transform, and only L of NF
// outputs are kept on each axis of the inverse, so both run one axis at a time on the lines that
// matter, in one buffer B of L * NF * NF complex values laid out (i0 < L, k1 < NF, k2 < NF):
//   Z+ : z lines of x (i0, i1 < L)               -> B[i0][i1][k2]       (L^2 lines, L of NF in)
//   Y+ : y lines at (i0, k2), rows i1 < L        -> B[i0][k1][k2]       (in place)
//   X  : x lines at (k1, k2), rows i0 < L: FFT, times lhs, inverse FFT, rows o < L back in place
//   Y- : y lines at (i0, k2), all NF rows        -> B[i0][o < L][k2]    (in place)
//   Z- : z lines (i0, i1 < L), all NF            -> y[i0][i1][i2 < L] / NF^3
// Every element a pass reads was written earlier in the same call, so B never needs clearing.
// Each workgroup owns whole lines (16 of them), reads all of its inputs before the FFT and writes
// after it, so the in-place passes are race-free.  The strided passes take 16 consecutive k2
// columns, so every row of a tile is one 128-byte segment.
//
// HBM traffic per apply at L = 128, NF = 256: 16.8 + 2 x 33.5 + 2 x 67 + 67 (+ 134 lhs) + ... =
// 0.57 GB, against ~2 GB for pad + full C2C FFT + multiply + full inverse + crop.
#include <rocfft/rocfft.h>

#include <map>
#include <mutex>
#include <tuple>

#include "fft_lds.h"

// Workspace of cov3d_toeplitz_apply_noalloc: the L * nf^2 complex buffer of the pruned passes for
// L * nf^2 <= 2^23 (64 MiB; L = 128 with nf = 256 fits).  Each device the library runs on holds
// its own copy from the first use until the process ends.  Host code never passes this name to a
// kernel: it uses the device address returned by hipGetSymbolAddress (cov3d::scratch_address).
constexpr int64_t kTplScratchElems = (int64_t)1 << 23;
static __device__ __align__(256) float2 scratch[kTplScratchElems];

namespace cov3d {

// ---- Z+ and Z-: contiguous z lines, 16 per workgroup.  g = i0 * L + i1 < L^2.
// FWD: x[g * L + i2] at position wrap(i2) -> B[(i0 NF + i1) NF + k2] for all k2.
// !FWD: B[(i0 NF + i1) NF + k2] for all k2 -> y[g * L + i2] = line[wrap(i2)] * scale.
template <int NF, bool FWD>
__global__ void __launch_bounds__(kFftThreads)
    k_tpl_z(const float2* __restrict__ in, float2* __restrict__ out, int L, float scale) {
  constexpr int LS = NF + 1;
  __shared__ float2 s_buf[kFftLines * LS];
  __shared__ float2 s_tw[NF];
  make_twiddles<NF>(s_tw);
  const int64_t nlines = (int64_t)L * L;
  const int64_t g0 = (int64_t)blockIdx.x * kFftLines;
  if constexpr (FWD) {
    for (int e = threadIdx.x; e < kFftLines * NF; e += kFftThreads) s_buf[(e / NF) * LS + e % NF] = make_float2(0.f, 0.f);
    __syncthreads();
    // the source of element e is in[g0 L + e]: kColGroup loads in flight per thread, the index
    // clamped to the last element so that they need no branch, then the stores of the valid ones
    const int64_t last = nlines * L - 1;
    for (int e0 = threadIdx.x; e0 < kFftLines * L; e0 += kFftThreads * kColGroup) {
      float2 v[kColGroup];
#pragma unroll
      for (int u = 0; u < kColGroup; ++u) {
        const int64_t src = g0 * L + e0 + u * kFftThreads;
        v[u] = in[src < last ? src : last];
      }
#pragma unroll
      for (int u = 0; u < kColGroup; ++u) {
        const int e = e0 + u * kFftThreads, l = e / L, i2 = e % L;
        if (e < kFftLines * L && g0 + l < nlines) s_buf[l * LS + wrap_mode(i2, L, NF)] = v[u];
      }
    }
  } else {
    // GB loads in flight per thread (kFftLines * NF / kFftThreads elements each, a power of two);
    // lines past the end read the last line
    constexpr int PER = kFftLines * NF / kFftThreads;
    constexpr int GB = PER < kColGroup ? PER : kColGroup;
    static_assert(PER % GB == 0 && PER * kFftThreads == kFftLines * NF, "whole batches of loads");
    for (int e0 = threadIdx.x; e0 < kFftLines * NF; e0 += kFftThreads * GB) {
      float2 v[GB];
#pragma unroll
      for (int u = 0; u < GB; ++u) {
        const int e = e0 + u * kFftThreads, l = e / NF, k = e % NF;
        const int64_t g = g0 + l < nlines ? g0 + l : nlines - 1;
        v[u] = in[((g / L) * NF + g % L) * NF + k];
      }
#pragma unroll
      for (int u = 0; u < GB; ++u) {
        const int e = e0 + u * kFftThreads;
        if (g0 + e / NF < nlines) s_buf[(e / NF) * LS + e % NF] = v[u];
      }
    }
  }
  __syncthreads();
  lds_fft<NF, LS>(s_buf, s_tw, FWD ? -1.f : 1.f);
  if constexpr (FWD) {
    for (int e = threadIdx.x; e < kFftLines * NF; e += kFftThreads) {
      const int l = e / NF, k = e % NF;
      const int64_t g = g0 + l;
      if (g < nlines) {
        const int64_t i0 = g / L, i1 = g % L;
        out[(i0 * NF + i1) * NF + k] = s_buf[l * LS + k];
      }
    }
  } else {
    for (int e = threadIdx.x; e < kFftLines * L; e += kFftThreads) {
      const int l = e / L, i2 = e % L;
      if (g0 + l < nlines) {
        const float2 v = s_buf[l * LS + wrap_mode(i2, L, NF)];
        out[(g0 + l) * L + i2] = make_float2(v.x * scale, v.y * scale);
      }
    }
  }
}

// ---- strided passes on B (i0 < L, k1 < NF, k2 < NF), 16 consecutive k2 columns per workgroup.
// MODE 0 (Y+): position (i0, c0); rows i1 < L at wrap(i1) -> FFT -1 -> all NF rows.
// MODE 1 (X):  position (k1, c0); rows i0 < L at wrap(i0) -> FFT -1 -> * lhs -> FFT +1 -> rows o < L.
// MODE 2 (Y-): position (i0, c0); all NF rows -> FFT +1 -> rows o < L from wrap(o).
template <int NF, int MODE>
__global__ void __launch_bounds__(kFftThreads)
    k_tpl_cols(float2* __restrict__ B, const float2* __restrict__ lhs, int L) {
  constexpr int LS = NF + 1;
  constexpr int NCT = NF / kFftLines;
  constexpr int RS = kFftThreads / kFftLines;  // rows per sweep of the 256 threads
  __shared__ float2 s_buf[kFftLines * LS];
  __shared__ float2 s_tw[NF];
  make_twiddles<NF>(s_tw);
  // a position's column tiles and neighbouring positions on one XCD (shared L2 lines)
  const int64_t lb = xcd_contiguous(blockIdx.x, gridDim.x);
  const int c0 = (int)(lb % NCT) * kFftLines;
  const int64_t pos = lb / NCT;
  const int col = threadIdx.x % kFftLines, rsub = threadIdx.x / kFftLines;
  // row r of this workgroup's lines sits at B[base + r * stride + c0 + col]
  const int64_t stride = MODE == 1 ? (int64_t)NF * NF : NF;
  const int64_t base = MODE == 1 ? pos * NF : pos * NF * NF;
  const bool pruned_in = MODE != 2;
  if (pruned_in) {
    for (int e = threadIdx.x; e < kFftLines * NF; e += kFftThreads) s_buf[(e / NF) * LS + e % NF] = make_float2(0.f, 0.f);
    __syncthreads();
  }
  const int nin = pruned_in ? L : NF;
  // kColGroup rows' loads in flight before their LDS stores (a store right after each load would
  // wait for it)
  for (int r0 = rsub; r0 < nin; r0 += RS * kColGroup) {
    float2 v[kColGroup];
#pragma unroll
    for (int u = 0; u < kColGroup; ++u) {
      const int r = r0 + u * RS;
      if (r < nin) v[u] = B[base + r * stride + c0 + col];
    }
#pragma unroll
    for (int u = 0; u < kColGroup; ++u) {
      const int r = r0 + u * RS;
      if (r < nin) s_buf[col * LS + (pruned_in ? wrap_mode(r, L, NF) : r)] = v[u];
    }
  }
  __syncthreads();
  lds_fft<NF, LS>(s_buf, s_tw, MODE == 2 ? 1.f : -1.f);
  if constexpr (MODE == 1) {
    // frequency (k0, k1 = pos, k2 = c0 + col) takes lhs at the centred index (k + NF/2) mod NF
    const int64_t c1 = (pos + NF / 2) % NF;
    const int c2 = (c0 + col + NF / 2) % NF;
    for (int k00 = rsub; k00 < NF; k00 += RS * kColGroup) {
      float2 bv[kColGroup];
#pragma unroll
      for (int u = 0; u < kColGroup; ++u) {
        const int k0 = k00 + u * RS;
        if (k0 < NF) bv[u] = lhs[((int64_t)((k0 + NF / 2) % NF) * NF + c1) * NF + c2];
      }
#pragma unroll
      for (int u = 0; u < kColGroup; ++u) {
        const int k0 = k00 + u * RS;
        if (k0 < NF) {
          const float2 a = s_buf[col * LS + k0], b = bv[u];
          s_buf[col * LS + k0] = make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
        }
      }
    }
    __syncthreads();
    lds_fft<NF, LS>(s_buf, s_tw, 1.f);
  }
  if constexpr (MODE == 0) {
    for (int r = rsub; r < NF; r += RS) B[base + r * stride + c0 + col] = s_buf[col * LS + r];
  } else {
    for (int o = rsub; o < L; o += RS) B[base + o * stride + c0 + col] = s_buf[col * LS + wrap_mode(o, L, NF)];
  }
}

template <int NF>
static int tpl_pruned_nf(const float2* x, const float2* lhs, float2* y, float2* B, int L, hipStream_t st) {
  constexpr int NCT = NF / kFftLines;
  const int64_t zl = ceil_div((int64_t)L * L, kFftLines);
  const float scale = (float)(1.0 / ((double)NF * NF * NF));
  hipLaunchKernelGGL((k_tpl_z<NF, true>), dim3((unsigned)zl), dim3(kFftThreads), 0, st, x, B, L, 1.f);
  COV3D_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_tpl_cols<NF, 0>), dim3((unsigned)((int64_t)L * NCT)), dim3(kFftThreads), 0, st, B, lhs, L);
  COV3D_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_tpl_cols<NF, 1>), dim3((unsigned)((int64_t)NF * NCT)), dim3(kFftThreads), 0, st, B, lhs, L);
  COV3D_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_tpl_cols<NF, 2>), dim3((unsigned)((int64_t)L * NCT)), dim3(kFftThreads), 0, st, B, lhs, L);
  COV3D_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_tpl_z<NF, false>), dim3((unsigned)zl), dim3(kFftThreads), 0, st, B, y, L, scale);
  COV3D_LAUNCH_CHECK();
  return 0;
}

bool tpl_pruned_supported(int64_t L, int64_t nf) {
  return (nf == 64 || nf == 128 || nf == 256 || nf == 512) && L >= 1 && L <= nf;
}

// y = crop(ifft3c(lhs * fft3c(pad(x)))) through the pruned passes, B: L * nf * nf complex values.
int tpl_pruned(const void* x, const void* lhs, void* y, void* B, int64_t L, int64_t nf, hipStream_t st) {
  const float2* xi = (const float2*)x;
  const float2* li = (const float2*)lhs;
  float2* yo = (float2*)y;
  float2* b = (float2*)B;
  switch (nf) {
    case 64: return tpl_pruned_nf<64>(xi, li, yo, b, (int)L, st);
    case 128: return tpl_pruned_nf<128>(xi, li, yo, b, (int)L, st);
    case 256: return tpl_pruned_nf<256>(xi, li, yo, b, (int)L, st);
    case 512: return tpl_pruned_nf<512>(xi, li, yo, b, (int)L, st);
    default: set_error("Toeplitz: unsupported grid %lld", (long long)nf); return COV3D_ERR_ARG;
  }
}

// ----------------------------------------------------------------------------
// Other grids: pad + rocFFT C2C + multiply + rocFFT inverse + crop on the full nf^3 grid.
// ----------------------------------------------------------------------------
__global__ void k_tpl_pad(const float2* __restrict__ x, float2* __restrict__ u, int L, int nf) {
  const int64_t n3 = (int64_t)nf * nf * nf;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n3) return;
  int j2 = (int)(idx % nf), j1 = (int)((idx / nf) % nf), j0 = (int)(idx / ((int64_t)nf * nf));
  auto src = [&](int j) -> int {
    int s = j < nf - L / 2 ? j : j - nf;  // signed offset
    int i = s + L / 2;
    return (i >= 0 && i < L) ? i : -1;
  };
  int i0 = src(j0), i1 = src(j1), i2 = src(j2);
  float2 v = make_float2(0.f, 0.f);
  if (i0 >= 0 && i1 >= 0 && i2 >= 0) v = x[((int64_t)i0 * L + i1) * L + i2];
  u[idx] = v;
}

__global__ void k_tpl_mult(float2* __restrict__ u, const float2* __restrict__ lhs, int nf) {
  const int64_t n3 = (int64_t)nf * nf * nf;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n3) return;
  int k2 = (int)(idx % nf), k1 = (int)((idx / nf) % nf), k0 = (int)(idx / ((int64_t)nf * nf));
  int c0 = (k0 + nf / 2) % nf, c1 = (k1 + nf / 2) % nf, c2 = (k2 + nf / 2) % nf;
  float2 a = u[idx], b = lhs[((int64_t)c0 * nf + c1) * nf + c2];
  u[idx] = make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__global__ void k_tpl_crop(const float2* __restrict__ v, float2* __restrict__ y, int L, int nf, float scale) {
  const int64_t n3 = (int64_t)L * L * L;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n3) return;
  int i2 = (int)(idx % L), i1 = (int)((idx / L) % L), i0 = (int)(idx / ((int64_t)L * L));
  auto dst = [&](int i) -> int {
    int s = i - L / 2;
    return s >= 0 ? s : s + nf;
  };
  float2 a = v[((int64_t)dst(i0) * nf + dst(i1)) * nf + dst(i2)];
  y[idx] = make_float2(a.x * scale, a.y * scale);
}

struct TplCtx {
  rocfft_plan fwd = nullptr, bwd = nullptr;
  rocfft_execution_info info = nullptr;
  void* work = nullptr;
  void* buf = nullptr;
  bool pruned = false;
};

static std::mutex g_tpl_mu;
static std::map<std::tuple<int, int64_t, int64_t>, TplCtx*> g_tpl;
static std::once_flag g_rocfft_once_tpl;

// The per-(device, L, nf) workspace of cov3d_toeplitz_apply, created on first use and kept.
static int tpl_get(int64_t L, int64_t nf, TplCtx** out) {
  int dev = 0;
  COV3D_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_tpl_mu);
  auto key = std::make_tuple(dev, L, nf);
  auto it = g_tpl.find(key);
  if (it != g_tpl.end()) {
    *out = it->second;
    return 0;
  }
  TplCtx* c = new TplCtx();
  if (tpl_pruned_supported(L, nf)) {
    c->pruned = true;
    if (hipMalloc(&c->buf, (size_t)L * nf * nf * sizeof(float2)) != hipSuccess) {
      delete c;
      set_error("hipMalloc of the Toeplitz buffer failed");
      return COV3D_ERR_HIP;
    }
    g_tpl[key] = c;
    *out = c;
    return 0;
  }
  std::call_once(g_rocfft_once_tpl, []() { rocfft_setup(); });
  size_t len[3] = {(size_t)nf, (size_t)nf, (size_t)nf};
  if (rocfft_plan_create(&c->fwd, rocfft_placement_inplace, rocfft_transform_type_complex_forward,
                         rocfft_precision_single, 3, len, 1, nullptr) != rocfft_status_success ||
      rocfft_plan_create(&c->bwd, rocfft_placement_inplace, rocfft_transform_type_complex_inverse,
                         rocfft_precision_single, 3, len, 1, nullptr) != rocfft_status_success ||
      rocfft_execution_info_create(&c->info) != rocfft_status_success) {
    set_error("rocfft plan creation failed for Toeplitz operator nf=%lld", (long long)nf);
    delete c;
    return COV3D_ERR_FFT;
  }
  size_t w1 = 0, w2 = 0;
  rocfft_plan_get_work_buffer_size(c->fwd, &w1);
  rocfft_plan_get_work_buffer_size(c->bwd, &w2);
  size_t ws = w1 > w2 ? w1 : w2;
  if (ws) {
    if (hipMalloc(&c->work, ws) != hipSuccess) {
      set_error("hipMalloc of rocFFT work buffer failed");
      return COV3D_ERR_HIP;
    }
    rocfft_execution_info_set_work_buffer(c->info, c->work, ws);
  }
  if (hipMalloc(&c->buf, (size_t)nf * nf * nf * 8) != hipSuccess) {
    set_error("hipMalloc of Toeplitz buffer failed");
    return COV3D_ERR_HIP;
  }
  g_tpl[key] = c;
  *out = c;
  return 0;
}

// Device address of `scratch` on the current device, looked up once per device with that device
// current and kept.
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_addr[kMaxDevices] = {};

static int scratch_address(void** out) {
  int dev = 0;
  COV3D_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device index %d out of range for the no-allocation workspace", dev);
    return COV3D_ERR_HIP;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error("hipGetSymbolAddress of the Toeplitz workspace failed: %s", hipGetErrorString(e));
      return COV3D_ERR_HIP;
    }
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return 0;
}

}  // namespace cov3d

extern "C" int cov3d_toeplitz_apply_noalloc(const void* x, const void* lhs, void* y, int64_t L, int64_t nf,
                                            void* stream) {
  COV3D_CHECK_ARG(L >= 1 && nf >= L, "need nf >= L >= 1");
  COV3D_CHECK_ARG(x && lhs && y, "NULL pointer");
  COV3D_CHECK_ARG(cov3d::tpl_pruned_supported(L, nf),
                  "no-allocation Toeplitz operator: nf must be 64, 128, 256 or 512 (got %lld)", (long long)nf);
  COV3D_CHECK_ARG(L * nf * nf <= kTplScratchElems,
                  "no-allocation Toeplitz operator: L * nf^2 = %lld exceeds the workspace of %lld elements",
                  (long long)(L * nf * nf), (long long)kTplScratchElems);
  void* buf = nullptr;
  int rc = cov3d::scratch_address(&buf);
  if (rc) return rc;
  return cov3d::tpl_pruned(x, lhs, y, buf, L, nf, cov3d::as_stream(stream));
}

extern "C" int cov3d_toeplitz_apply(const void* x, const void* lhs, void* y, int64_t L, int64_t nf,
                                    void* stream) {
  COV3D_CHECK_ARG(L >= 1 && nf >= L, "need nf >= L >= 1");
  COV3D_CHECK_ARG(x && lhs && y, "NULL pointer");
  cov3d::TplCtx* c = nullptr;
  int rc = cov3d::tpl_get(L, nf, &c);
  if (rc) return rc;
  hipStream_t st = cov3d::as_stream(stream);
  if (c->pruned) return cov3d::tpl_pruned(x, lhs, y, c->buf, L, nf, st);
  const int64_t n3 = nf * nf * nf, l3 = L * L * L;
  const int blk = 256;
  hipLaunchKernelGGL(cov3d::k_tpl_pad, dim3((unsigned)cov3d::ceil_div(n3, blk)), dim3(blk), 0, st,
                     (const float2*)x, (float2*)c->buf, (int)L, (int)nf);
  COV3D_LAUNCH_CHECK();
  void* bufs[1] = {c->buf};
  if (rocfft_execution_info_set_stream(c->info, st) != rocfft_status_success ||
      rocfft_execute(c->fwd, bufs, nullptr, c->info) != rocfft_status_success) {
    cov3d::set_error("rocfft forward failed in Toeplitz operator");
    return COV3D_ERR_FFT;
  }
  hipLaunchKernelGGL(cov3d::k_tpl_mult, dim3((unsigned)cov3d::ceil_div(n3, blk)), dim3(blk), 0, st,
                     (float2*)c->buf, (const float2*)lhs, (int)nf);
  COV3D_LAUNCH_CHECK();
  if (rocfft_execute(c->bwd, bufs, nullptr, c->info) != rocfft_status_success) {
    cov3d::set_error("rocfft inverse failed in Toeplitz operator");
    return COV3D_ERR_FFT;
  }
  hipLaunchKernelGGL(cov3d::k_tpl_crop, dim3((unsigned)cov3d::ceil_div(l3, blk)), dim3(blk), 0, st,
                     (const float2*)c->buf, (float2*)y, (int)L, (int)nf, (float)(1.0 / (double)n3));
  COV3D_LAUNCH_CHECK();
  return 0;
}
