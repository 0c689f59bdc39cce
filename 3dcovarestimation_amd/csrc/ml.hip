// Maximum-likelihood contractions of rank-r projected eigenvectors.
//
// Replaces compute_latentMAP_batch (cov3d/wiener_coords.py:168-203) and the
// Gram / solve / logdet / autograd work of cost_maximum_liklihood
// (cov3d/covar_sgd.py:1238-1286).
//
// One workgroup per image.  The rows [Q_0 .. Q_{r-1}, y~] (complex, d pixels,
// treated as 2d real components) are read once from HBM; the (r+1)x(r+1)
// real Gram is accumulated either with FP32 FMAs (precision 0) or with
// v_mfma_f32_32x32x16_bf16 on split-BF16 operands v = hi + lo, summing
// hi*hi + hi*lo + lo*hi in FP32 (precision 1).  The r x r system is then
// factorised in FP64 (Cholesky) inside the same workgroup.
#include <math.h>

#include "common.h"

namespace cov3d {

constexpr int kMlBlock = 256;
constexpr int kRmax = 32;  // rows of the Gram tile (r + 1 <= 32)
// Split-BF16 forward: 8 waves per image.  Each wave takes 32-pixel chunks: its lanes load the
// chunk's rows coalesced (lane = pixel, two rows per instruction), scale and split them, and
// store the BF16 pairs in a per-wave LDS tile [row][pixel], from which the MFMA operands are
// read as 16-byte rows.
constexpr int kMlBlockMfma = 512;
constexpr int kMlChunk = 32;
constexpr int kMlTileLd = kMlChunk + 4;  // u32 per tile row (the pad spreads rows over banks)
// LDS of the reduction / solve phase: s_part[4][32 x 32] f32, then s_G, s_A, s_minv ((32 x 33) f64
// each) and s_z (32 f64)
constexpr int kMlPartBytes = 4 * kRmax * kRmax * 4;
constexpr int kMlPhase2Bytes = kMlPartBytes + 3 * kRmax * (kRmax + 1) * 8 + kRmax * 8;
constexpr int kMlTileBytes = (kMlBlockMfma / 64) * 2 * kRmax * kMlTileLd * 4;
constexpr int kMlRowsBytes = kRmax * 64 * 8;  // FP32 path: 64 pixels of every row

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ unsigned short f32_to_bf16_bits(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_bits_to_f32(unsigned short h) {
  return __uint_as_float(((unsigned)h) << 16);
}

// Scale of pixel k of image i: mask * ctf / L^2 (mask zeroes row/col 0 for even L).
__device__ __forceinline__ float pixel_scale(const float* __restrict__ ctf, int64_t i, int64_t k,
                                             int L, float inv_l2) {
  if ((L & 1) == 0) {
    int ky = (int)(k / L), kx = (int)(k - (int64_t)ky * L);
    if (ky == 0 || kx == 0) return 0.f;
  }
  float s = inv_l2;
  if (ctf) s *= ctf[i * (int64_t)L * L + k];
  return s;
}

// Pixel of entry k of image i's point list: the identity (full plane) or a
// map onto the centred L x L grid (Hermitian half plane); mult is the number
// of grid pixels the entry stands for (2 for a +-k pair, 1 for the centre).
__device__ __forceinline__ int64_t entry_pixel(const int* __restrict__ pix_map, int64_t k, int L,
                                               float& mult) {
  if (!pix_map) {
    mult = 1.f;
    return k;
  }
  const int64_t px = pix_map[k];
  mult = px == (int64_t)(L / 2) * L + L / 2 ? 1.f : 2.f;
  return px;
}

// Row value (complex) of entry k for Gram row m of image i, scaled by
// sqrt(mult) so that the plain Gram of the rows is the full-plane sum.
__device__ __forceinline__ float2 gram_row_value(const float2* __restrict__ c,
                                                 const float2* __restrict__ yhat,
                                                 const float* __restrict__ ctf,
                                                 const int* __restrict__ pix_map, int64_t i, int m,
                                                 int r, int64_t B, int64_t np, int64_t k, int L,
                                                 float inv_l, float inv_l2) {
  if (k >= np || m > r) return make_float2(0.f, 0.f);
  float mult;
  const int64_t px = entry_pixel(pix_map, k, L, mult);
  const float sm = mult == 1.f ? 1.f : 1.41421356237309515f;
  const int64_t d = (int64_t)L * L;
  if (m == r) {
    float2 v = yhat[i * d + px];
    return make_float2(v.x * inv_l * sm, v.y * inv_l * sm);
  }
  float s = pixel_scale(ctf, i, px, L, inv_l2) * sm;
  float2 v = c[((int64_t)m * B + i) * np + k];
  return make_float2(v.x * s, v.y * s);
}

// FP64 Cholesky of the r x r matrix in A (leading dimension kRmax+1), in place
// (lower triangle).  All threads of the block take part.
__device__ void block_cholesky(double* A, int r) {
  const int tid = threadIdx.x;
  for (int k = 0; k < r; ++k) {
    if (tid == 0) A[k * (kRmax + 1) + k] = sqrt(A[k * (kRmax + 1) + k]);
    __syncthreads();
    const double dkk = A[k * (kRmax + 1) + k];
    for (int i = k + 1 + tid; i < r; i += blockDim.x) A[i * (kRmax + 1) + k] /= dkk;
    __syncthreads();
    const int n = r - k - 1;
    for (int q = tid; q < n * n; q += blockDim.x) {
      int i = k + 1 + q / n, j = k + 1 + q % n;
      if (j <= i) A[i * (kRmax + 1) + j] -= A[i * (kRmax + 1) + k] * A[j * (kRmax + 1) + k];
    }
    __syncthreads();
  }
}

// Solves (L L^T) x = b in place for one right-hand side (serial, r <= 31).
__device__ void chol_solve(const double* A, double* x, int r) {
  for (int i = 0; i < r; ++i) {
    double s = x[i];
    for (int k = 0; k < i; ++k) s -= A[i * (kRmax + 1) + k] * x[k];
    x[i] = s / A[i * (kRmax + 1) + i];
  }
  for (int i = r - 1; i >= 0; --i) {
    double s = x[i];
    for (int k = i + 1; k < r; ++k) s -= A[k * (kRmax + 1) + i] * x[k];
    x[i] = s / A[i * (kRmax + 1) + i];
  }
}

template <int PREC, int NR2 = 16>
__global__ void __launch_bounds__(PREC == 1 ? kMlBlockMfma : kMlBlock)
    k_ml_forward(const float2* __restrict__ c, const float* __restrict__ ctf,
                 const float2* __restrict__ yhat, const float* __restrict__ lam_inv,
                 const int* __restrict__ pix_map, int64_t np, int64_t B,
                 int r, int L, float noise_var, int mean_const, float* __restrict__ cost,
                 float* __restrict__ z_out, float* __restrict__ m_out,
                 float* __restrict__ minv_out, float* __restrict__ b_out) {
  // The split-BF16 tiles and the reduction arrays are never live together: one buffer holds both.
  constexpr int kSmem = PREC == 1 ? (kMlTileBytes > kMlPhase2Bytes ? kMlTileBytes : kMlPhase2Bytes)
                                  : kMlPhase2Bytes + kMlRowsBytes;
  __shared__ __align__(16) unsigned char s_raw[kSmem];
  float(*s_part)[kRmax * kRmax] = reinterpret_cast<float(*)[kRmax * kRmax]>(s_raw);
  double* s_G = reinterpret_cast<double*>(s_raw + kMlPartBytes);
  double* s_A = s_G + kRmax * (kRmax + 1);
  double* s_minv = s_A + kRmax * (kRmax + 1);
  double* s_z = s_minv + kRmax * (kRmax + 1);
  const int64_t i = blockIdx.x;
  const int64_t d = (int64_t)L * L;
  const float inv_l = 1.f / (float)L;
  const float inv_l2 = inv_l * inv_l;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;

  if constexpr (PREC == 1) {
    // ---- split-BF16 MFMA Gram: wave w of nw covers 32-pixel chunks w, w + nw, ...
    const int nw = (int)(blockDim.x >> 6);
    unsigned* t_hi = reinterpret_cast<unsigned*>(s_raw) + wave * 2 * kRmax * kMlTileLd;
    unsigned* t_lo = t_hi + kRmax * kMlTileLd;
    // rows above r stay zero
    for (int e = (r + 1) * kMlTileLd + lane; e < kRmax * kMlTileLd; e += 64) t_hi[e] = t_lo[e] = 0u;
    const int p = lane & 31, rp = lane >> 5;  // load role: pixel p, rows rp, rp + 2, ...
    const int m = lane & 31, h = lane >> 5;   // MFMA role: row m, pixel quarter h
    const int nrow2 = (r + 2) >> 1;           // row pairs covering rows 0..r (<= NR2)
    f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.f;
    const int64_t nchunk = (np + kMlChunk - 1) / kMlChunk;
    // loads of one chunk: row rp + 2j of pixel p (the image row r comes from yhat), the pixel's
    // scale; addresses are clamped so that every lane issues every load (a fixed count of
    // loads in flight), and out-of-range values are zeroed when the chunk is stored
    auto load = [&](int64_t ch, float2 (&v)[NR2], float& sc, float& sy) {
      const int64_t k = ch * kMlChunk + p;
      const int64_t kc = k < np ? k : np - 1;
      float mult;
      const int64_t px = entry_pixel(pix_map, kc, L, mult);
      const float sm = mult == 1.f ? 1.f : 1.41421356237309515f;
      sc = k < np ? pixel_scale(ctf, i, px, L, inv_l2) * sm : 0.f;
      sy = k < np ? inv_l * sm : 0.f;
#pragma unroll
      for (int j = 0; j < NR2; ++j) {
        const int mm = rp + 2 * j;
        const float2* src = mm == r ? yhat + (i * d + px) : c + (((int64_t)(mm < r ? mm : 0) * B + i) * np + kc);
        v[j] = *src;
      }
    };
    auto store_mma = [&](const float2 (&v)[NR2], float sc, float sy) {
#pragma unroll
      for (int j = 0; j < NR2; ++j) {
        const int mm = rp + 2 * j;
        if (j < nrow2 && mm <= r) {
          const float sv = mm == r ? sy : sc;
          const float x = v[j].x * sv, y = v[j].y * sv;
          const unsigned short hx = f32_to_bf16_bits(x), hy = f32_to_bf16_bits(y);
          const unsigned short lx = f32_to_bf16_bits(x - bf16_bits_to_f32(hx));
          const unsigned short ly = f32_to_bf16_bits(y - bf16_bits_to_f32(hy));
          t_hi[mm * kMlTileLd + p] = (unsigned)hx | ((unsigned)hy << 16);
          t_lo[mm * kMlTileLd + p] = (unsigned)lx | ((unsigned)ly << 16);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int q = 0; q < kMlChunk / 8; ++q) {
        const uint4 uh = *reinterpret_cast<const uint4*>(&t_hi[m * kMlTileLd + 8 * q + 4 * h]);
        const uint4 ul = *reinterpret_cast<const uint4*>(&t_lo[m * kMlTileLd + 8 * q + 4 * h]);
        const bf16x8 ah = __builtin_bit_cast(bf16x8, uh);
        const bf16x8 al = __builtin_bit_cast(bf16x8, ul);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, ah, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, al, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, ah, acc, 0, 0, 0);
      }
      // the next chunk's stores must not overtake this chunk's operand reads
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    // two register sets: the loads of chunk ch + nw are in flight while chunk ch is stored and
    // multiplied (the last prefetch repeats a valid chunk and is not used)
    if (wave < nchunk) {
      float2 va[NR2], vb[NR2];
      float sca, sya, scb, syb;
      load(wave, va, sca, sya);
      for (int64_t ch = wave; ch < nchunk; ch += 2 * nw) {
        const int64_t c1 = ch + nw < nchunk ? ch + nw : ch;
        load(c1, vb, scb, syb);
        store_mma(va, sca, sya);
        if (ch + nw >= nchunk) break;
        const int64_t c2 = ch + 2 * nw < nchunk ? ch + 2 * nw : ch + nw;
        load(c2, va, sca, sya);
        store_mma(vb, scb, syb);
      }
    }
    __syncthreads();  // the tiles are dead: s_part aliases them
    // D layout of the 32x32 tile: lane l, register v -> row 8*(v/4) + 4*(l/32) + v%4, col l%32.
    // Waves are folded into the 4 partial tiles in rounds: round j adds waves 4j .. 4j + 3.
    for (int j = 0; 4 * j < nw; ++j) {
      if ((wave >> 2) == j) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          int row = 8 * (v >> 2) + 4 * h + (v & 3);
          float* dst = &s_part[wave & 3][row * kRmax + m];
          *dst = (j == 0 ? 0.f : *dst) + acc[v];
        }
      }
      __syncthreads();
    }
  } else {
    // ---- FP32 Gram (256 threads): stage 64 pixels of all rows in LDS; thread k owns the
    // upper-triangle pairs k, k + 256, k + 512 ((r+1)(r+2)/2 <= 528).
    float2(*s_rows)[64] = reinterpret_cast<float2(*)[64]>(s_raw + kMlPhase2Bytes);
    const int npair = (r + 1) * (r + 2) / 2;
    int pt[3], ps[3];
    float accp[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      int q = tid + k * kMlBlock, t = 0;
      pt[k] = ps[k] = -1;
      if (q < npair) {
        while (q >= (r + 1 - t)) {
          q -= (r + 1 - t);
          ++t;
        }
        pt[k] = t;
        ps[k] = t + q;
      }
    }
    for (int64_t k0 = 0; k0 < np; k0 += 64) {
      for (int e = tid; e < (r + 1) * 64; e += blockDim.x) {
        int m = e / 64, kk = e % 64;
        s_rows[m][kk] = gram_row_value(c, yhat, ctf, pix_map, i, m, r, B, np, k0 + kk, L, inv_l, inv_l2);
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (pt[k] >= 0) {
          for (int kk = 0; kk < 64; ++kk) {
            float2 a = s_rows[pt[k]][kk], bb = s_rows[ps[k]][kk];
            accp[k] += a.x * bb.x + a.y * bb.y;
          }
        }
      }
      __syncthreads();
    }
    for (int e = tid; e < 4 * kRmax * kRmax; e += blockDim.x) (&s_part[0][0])[e] = 0.f;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (pt[k] >= 0) {
        s_part[0][pt[k] * kRmax + ps[k]] = accp[k];
        s_part[0][ps[k] * kRmax + pt[k]] = accp[k];
      }
    }
    __syncthreads();
  }

  // ---- reduce partial Grams, form M, b
  for (int e = tid; e < (r + 1) * (r + 1); e += blockDim.x) {
    int a = e / (r + 1), b = e % (r + 1);
    double g = 0;
    for (int w = 0; w < 4; ++w) g += (double)s_part[w][a * kRmax + b];
    s_G[a * (kRmax + 1) + b] = g;
  }
  __syncthreads();
  const double inv_nv = 1.0 / (double)noise_var;
  for (int e = tid; e < r * r; e += blockDim.x) {
    int a = e / r, b = e % r;
    double g = 0.5 * (s_G[a * (kRmax + 1) + b] + s_G[b * (kRmax + 1) + a]) * inv_nv;
    g += lam_inv ? (double)lam_inv[a * r + b] : (a == b ? 1.0 : 0.0);
    s_A[a * (kRmax + 1) + b] = g;
    if (m_out) m_out[(i * r + a) * r + b] = (float)g;
  }
  __syncthreads();
  block_cholesky(s_A, r);
  // The triangular solves keep their vectors in LDS (s_z, and column tid of s_minv): private
  // arrays indexed by the runtime rank would live in scratch memory.
  if (tid == 0) {
    for (int a = 0; a < r; ++a) s_z[a] = s_G[a * (kRmax + 1) + r] * inv_nv;
    chol_solve(s_A, s_z, r);
    double bz = 0, logdet = 0;
    for (int a = 0; a < r; ++a) {
      const double ba = s_G[a * (kRmax + 1) + r] * inv_nv;
      bz += ba * s_z[a];
      logdet += 2.0 * log(s_A[a * (kRmax + 1) + a]);
      if (b_out) b_out[i * r + a] = (float)ba;
      if (z_out) z_out[i * r + a] = (float)s_z[a];
    }
    double cst = 0.5 * (-bz + logdet);
    if (mean_const) cst += 0.5 * s_G[r * (kRmax + 1) + r] * inv_nv;
    if (cost) cost[i] = (float)cst;
  }
  if (minv_out && tid < r) {
    double* x = s_minv + tid * (kRmax + 1);
    for (int a = 0; a < r; ++a) x[a] = (a == tid) ? 1.0 : 0.0;
    chol_solve(s_A, x, r);
    for (int a = 0; a < r; ++a) minv_out[(i * r + a) * r + tid] = (float)x[a];
  }
}

// Gradient images: g_t = mask*ctf*(Minv Q + z (Q^T z - y~))_t * scale, one thread per pixel.
// RM is a compile-time bound on r so the per-pixel rows stay in registers.
template <int RM>
__global__ void __launch_bounds__(kMlBlock)
    k_ml_grad(const float2* __restrict__ c, const float* __restrict__ ctf,
              const float2* __restrict__ yhat, const float* __restrict__ z,
              const float* __restrict__ minv, const int* __restrict__ pix_map, int64_t np,
              float2* __restrict__ g, int64_t B, int r, int L, float scale) {
  __shared__ float s_minv[RM * RM];
  __shared__ float s_z[RM];
  const int64_t i = blockIdx.y;
  const int64_t d = (int64_t)L * L;
  for (int e = threadIdx.x; e < RM * RM; e += blockDim.x) {
    int a = e / RM, b = e % RM;
    s_minv[e] = (a < r && b < r) ? minv[(i * r + a) * r + b] : 0.f;
  }
  if (threadIdx.x < RM) s_z[threadIdx.x] = (int)threadIdx.x < r ? z[i * r + threadIdx.x] : 0.f;
  __syncthreads();
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  float mult;
  const int64_t px = entry_pixel(pix_map, k, L, mult);
  const float inv_l = 1.f / (float)L;
  const float ps = pixel_scale(ctf, i, px, L, inv_l * inv_l);
  float qx[RM], qy[RM];
  float ex = 0.f, ey = 0.f;
#pragma unroll
  for (int s = 0; s < RM; ++s) {
    float2 v = s < r ? c[((int64_t)s * B + i) * np + k] : make_float2(0.f, 0.f);
    qx[s] = v.x * ps;
    qy[s] = v.y * ps;
    ex += s_z[s] * qx[s];
    ey += s_z[s] * qy[s];
  }
  float2 yv = yhat[i * d + px];
  ex -= yv.x * inv_l;
  ey -= yv.y * inv_l;
  const float gs = ps * scale * (float)(L * L) * mult;  // mult*mask*ctf*scale (ps carries 1/L^2)
  for (int t = 0; t < r; ++t) {
    float ax = s_z[t] * ex, ay = s_z[t] * ey;
#pragma unroll
    for (int s = 0; s < RM; ++s) {
      float mv = s_minv[t * RM + s];
      ax += mv * qx[s];
      ay += mv * qy[s];
    }
    g[((int64_t)t * B + i) * np + k] = make_float2(ax * gs, ay * gs);
  }
}

// Complex-coefficient gradient images of the Fourier-domain ML cost (complex M, z): with
// conj(M^-1) = R - iJ and z = a + ib the gradient splits into two real-coefficient passes
// (coefficients (A1, z1), (A2, z2) as the host forms them), g = g1 + i g2; both run here on one
// load of the pixel's r values, and g is written once.
template <int RM>
__global__ void __launch_bounds__(kMlBlock)
    k_ml_grad_cplx(const float2* __restrict__ c, const float* __restrict__ ctf,
                   const float2* __restrict__ yhat, const float* __restrict__ z1,
                   const float* __restrict__ m1, const float* __restrict__ z2, const float* __restrict__ m2,
                   const int* __restrict__ pix_map, int64_t np, float2* __restrict__ g, int64_t B, int r, int L,
                   float scale) {
  __shared__ float s_m1[RM * RM], s_m2[RM * RM];
  __shared__ float s_z1[RM], s_z2[RM];
  const int64_t i = blockIdx.y;
  const int64_t d = (int64_t)L * L;
  for (int e = threadIdx.x; e < RM * RM; e += blockDim.x) {
    int a = e / RM, b = e % RM;
    const bool in = a < r && b < r;
    s_m1[e] = in ? m1[(i * r + a) * r + b] : 0.f;
    s_m2[e] = in ? m2[(i * r + a) * r + b] : 0.f;
  }
  if (threadIdx.x < RM) {
    s_z1[threadIdx.x] = (int)threadIdx.x < r ? z1[i * r + threadIdx.x] : 0.f;
    s_z2[threadIdx.x] = (int)threadIdx.x < r ? z2[i * r + threadIdx.x] : 0.f;
  }
  __syncthreads();
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  float mult;
  const int64_t px = entry_pixel(pix_map, k, L, mult);
  const float inv_l = 1.f / (float)L;
  const float ps = pixel_scale(ctf, i, px, L, inv_l * inv_l);
  float qx[RM], qy[RM];
  float e1x = 0.f, e1y = 0.f, e2x = 0.f, e2y = 0.f;
#pragma unroll
  for (int s = 0; s < RM; ++s) {
    float2 v = s < r ? c[((int64_t)s * B + i) * np + k] : make_float2(0.f, 0.f);
    qx[s] = v.x * ps;
    qy[s] = v.y * ps;
    e1x += s_z1[s] * qx[s];
    e1y += s_z1[s] * qy[s];
    e2x += s_z2[s] * qx[s];
    e2y += s_z2[s] * qy[s];
  }
  const float2 yv = yhat[i * d + px];
  e1x -= yv.x * inv_l;
  e1y -= yv.y * inv_l;
  e2x -= yv.x * inv_l;
  e2y -= yv.y * inv_l;
  const float gs = ps * scale * (float)(L * L) * mult;
  for (int t = 0; t < r; ++t) {
    float ax = s_z1[t] * e1x, ay = s_z1[t] * e1y;
    float bx = s_z2[t] * e2x, by = s_z2[t] * e2y;
#pragma unroll
    for (int s = 0; s < RM; ++s) {
      const float u = s_m1[t * RM + s], w = s_m2[t * RM + s];
      ax += u * qx[s];
      ay += u * qy[s];
      bx += w * qx[s];
      by += w * qy[s];
    }
    g[((int64_t)t * B + i) * np + k] = make_float2((ax - by) * gs, (ay + bx) * gs);
  }
}

// Gradient images written straight into the sorted rows a type-1 NUFFT plan spreads from:
// value t of pixel k of image i goes to gs[spos[i * np + k] * ldc + t] (t < r; zeros up to ldc).
// Each workgroup stages its 256 pixels' rows in LDS and writes them as consecutive 16-byte pieces.
constexpr int kGradLdc = 24;  // largest row stride staged
template <int RM>
__global__ void __launch_bounds__(kMlBlock)
    k_ml_grad_sorted(const float2* __restrict__ c, const float* __restrict__ ctf,
                     const float2* __restrict__ yhat, const float* __restrict__ z,
                     const float* __restrict__ minv, const int* __restrict__ pix_map, int64_t np,
                     float2* __restrict__ gs, const int* __restrict__ spos, int ldc, int64_t B, int r, int L,
                     float scale) {
  __shared__ float s_minv[RM * RM];
  __shared__ float s_z[RM];
  __shared__ float2 s_v[kMlBlock * (kGradLdc + 1)];
  __shared__ int s_pos[kMlBlock];
  const int64_t i = blockIdx.y;
  const int64_t d = (int64_t)L * L;
  const int tid = threadIdx.x;
  for (int e = tid; e < RM * RM; e += blockDim.x) {
    int a = e / RM, b = e % RM;
    s_minv[e] = (a < r && b < r) ? minv[(i * r + a) * r + b] : 0.f;
  }
  if (tid < RM) s_z[tid] = tid < r ? z[i * r + tid] : 0.f;
  __syncthreads();
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + tid;
  float2* row = s_v + tid * (kGradLdc + 1);
  s_pos[tid] = k < np ? spos[i * np + k] : -1;
  if (r & 1) row[r] = make_float2(0.f, 0.f);
  if (k < np) {
    float mult;
    const int64_t px = entry_pixel(pix_map, k, L, mult);
    const float inv_l = 1.f / (float)L;
    const float ps = pixel_scale(ctf, i, px, L, inv_l * inv_l);
    float qx[RM], qy[RM];
    float ex = 0.f, ey = 0.f;
#pragma unroll
    for (int s2 = 0; s2 < RM; ++s2) {
      float2 v = s2 < r ? c[((int64_t)s2 * B + i) * np + k] : make_float2(0.f, 0.f);
      qx[s2] = v.x * ps;
      qy[s2] = v.y * ps;
      ex += s_z[s2] * qx[s2];
      ey += s_z[s2] * qy[s2];
    }
    float2 yv = yhat[i * d + px];
    ex -= yv.x * inv_l;
    ey -= yv.y * inv_l;
    const float gsc = ps * scale * (float)(L * L) * mult;
    for (int t = 0; t < r; ++t) {
      float ax = s_z[t] * ex, ay = s_z[t] * ey;
#pragma unroll
      for (int s2 = 0; s2 < RM; ++s2) {
        float mv = s_minv[t * RM + s2];
        ax += mv * qx[s2];
        ay += mv * qy[s2];
      }
      row[t] = make_float2(ax * gsc, ay * gsc);
    }
  }
  __syncthreads();
  // 16-byte pieces carrying the r values (the row padding up to ldc is never read)
  const int hp = (r + 1) >> 1;
  for (int e = tid; e < kMlBlock * hp; e += kMlBlock) {
    const int q = e / hp, h = e - q * hp;
    const int sp = s_pos[q];
    if (sp >= 0) {
      const float2 a = s_v[q * (kGradLdc + 1) + 2 * h], b = s_v[q * (kGradLdc + 1) + 2 * h + 1];
      reinterpret_cast<float4*>(gs + (int64_t)sp * ldc)[h] = float4{a.x, a.y, b.x, b.y};
    }
  }
}

// Batched Hermitian positive definite solve in FP64, one 64-lane workgroup per r x r matrix
// (r <= kHermMax): M = L L^H (Cholesky in LDS, column k's rows in parallel), log det M =
// 2 sum log L_kk, Y = L^-1 (one column per lane by forward substitution), M^-1 = Y^H Y and
// z = M^-1 b.  Complex values are (re, im) double pairs; the matrices are read and written
// row-major.  A non-positive pivot (M not positive definite) gives NaN outputs for that matrix.
constexpr int kHermMax = 32;
constexpr int kHermThreads = 64;

__global__ void __launch_bounds__(kHermThreads)
    k_herm_solve(const double2* __restrict__ M, const double2* __restrict__ b, double2* __restrict__ minv,
                 double2* __restrict__ z, double* __restrict__ logdet, int r) {
  __shared__ double2 s_L[kHermMax][kHermMax + 1];
  __shared__ double2 s_Y[kHermMax][kHermMax + 1];
  __shared__ double2 s_b[kHermMax];
  __shared__ double s_ld[kHermMax];
  const int64_t m = blockIdx.x;
  const int tid = threadIdx.x;
  const double2* Mm = M + m * r * r;
  for (int e = tid; e < r * r; e += kHermThreads) s_L[e / r][e % r] = Mm[e];
  if (tid < r) s_b[tid] = b[m * r + tid];
  __syncthreads();
  // right-looking Cholesky on the lower triangle: after step k, column k holds L(:, k)
  for (int k = 0; k < r; ++k) {
    if (tid == 0) s_L[k][k].x = sqrt(s_L[k][k].x), s_L[k][k].y = 0.0;
    __syncthreads();
    const double d = s_L[k][k].x;
    for (int i = k + 1 + tid; i < r; i += kHermThreads) {
      s_L[i][k].x /= d;
      s_L[i][k].y /= d;
    }
    __syncthreads();
    // trailing update A(i, j) -= L(i, k) conj(L(j, k)), k < j <= i
    for (int e = tid; e < (r - k - 1) * (r - k - 1); e += kHermThreads) {
      const int i = k + 1 + e / (r - k - 1), j = k + 1 + e % (r - k - 1);
      if (j <= i) {
        const double2 a = s_L[i][k], c = s_L[j][k];
        s_L[i][j].x -= a.x * c.x + a.y * c.y;
        s_L[i][j].y -= a.y * c.x - a.x * c.y;
      }
    }
    __syncthreads();
  }
  if (tid < r) s_ld[tid] = log(s_L[tid][tid].x);
  // Y = L^-1: lane c solves L y = e_c
  if (tid < r) {
    const int c = tid;
    for (int i = 0; i < r; ++i) {
      double2 acc = make_double2(i == c ? 1.0 : 0.0, 0.0);
      for (int j = c; j < i; ++j) {
        const double2 l = s_L[i][j], y = s_Y[j][c];
        acc.x -= l.x * y.x - l.y * y.y;
        acc.y -= l.x * y.y + l.y * y.x;
      }
      const double d = s_L[i][i].x;
      s_Y[i][c] = i < c ? make_double2(0.0, 0.0) : make_double2(acc.x / d, acc.y / d);
    }
  }
  __syncthreads();
  if (tid == 0) {
    double acc = 0.0;
    for (int k = 0; k < r; ++k) acc += s_ld[k];
    logdet[m] = 2.0 * acc;
  }
  // M^-1 = Y^H Y: element (i, j) = sum_k conj(Y(k, i)) Y(k, j), k >= max(i, j)
  double2* Im = minv + m * r * r;
  for (int e = tid; e < r * r; e += kHermThreads) {
    const int i = e / r, j = e % r;
    double2 acc = make_double2(0.0, 0.0);
    for (int k = max(i, j); k < r; ++k) {
      const double2 a = s_Y[k][i], c = s_Y[k][j];
      acc.x += a.x * c.x + a.y * c.y;
      acc.y += a.x * c.y - a.y * c.x;
    }
    Im[e] = acc;
    s_L[i][j] = acc;  // L is no longer needed: keep M^-1 for z
  }
  __syncthreads();
  if (tid < r) {
    double2 acc = make_double2(0.0, 0.0);
    for (int j = 0; j < r; ++j) {
      const double2 a = s_L[tid][j], c = s_b[j];
      acc.x += a.x * c.x - a.y * c.y;
      acc.y += a.x * c.y + a.y * c.x;
    }
    z[m * r + tid] = acc;
  }
}

}  // namespace cov3d

extern "C" int cov3d_ml_forward(const void* c, const void* ctf, const void* yhat,
                                const void* lam_inv, const int* pix_map, int64_t npix, int64_t B,
                                int64_t r, int64_t L, double noise_var, int mean_const,
                                int precision, void* cost, void* z, void* m_out, void* minv,
                                void* b_out, void* stream) {
  COV3D_CHECK_ARG(r >= 1 && r <= 31, "rank r=%lld must be in [1, 31]", (long long)r);
  COV3D_CHECK_ARG(B >= 0 && L >= 1, "bad batch/size");
  COV3D_CHECK_ARG(noise_var > 0, "noise_var must be positive");
  COV3D_CHECK_ARG(precision == 0 || precision == 1, "precision must be 0 or 1");
  if (B == 0) return 0;
  COV3D_CHECK_ARG(c && yhat, "c and yhat must not be NULL");
  COV3D_CHECK_ARG(pix_map ? (npix >= 1 && npix <= L * L) : (npix == 0 || npix == L * L),
                  "npix=%lld invalid for L=%lld", (long long)npix, (long long)L);
  const int64_t np = pix_map ? npix : L * L;
  hipStream_t st = cov3d::as_stream(stream);
  // split-BF16: register sets sized for the (r + 2) / 2 row pairs a lane loads
#define COV3D_MLF(NR)                                                                                  \
  hipLaunchKernelGGL((cov3d::k_ml_forward<1, NR>), dim3((unsigned)B), dim3(cov3d::kMlBlockMfma), 0, st, \
                     (const float2*)c, (const float*)ctf, (const float2*)yhat, (const float*)lam_inv, pix_map, \
                     np, B, (int)r, (int)L, (float)noise_var, mean_const, (float*)cost, (float*)z,          \
                     (float*)m_out, (float*)minv, (float*)b_out)
  const int64_t nr2 = (r + 2) / 2;
  if (precision == 1 && nr2 <= 4)
    COV3D_MLF(4);
  else if (precision == 1 && nr2 <= 8)
    COV3D_MLF(8);
  else if (precision == 1 && nr2 <= 11)
    COV3D_MLF(11);
  else if (precision == 1)
    COV3D_MLF(16);
  else
    hipLaunchKernelGGL((cov3d::k_ml_forward<0>), dim3((unsigned)B), dim3(cov3d::kMlBlock), 0, st,
                       (const float2*)c, (const float*)ctf, (const float2*)yhat,
                       (const float*)lam_inv, pix_map, np, B, (int)r, (int)L, (float)noise_var, mean_const,
                       (float*)cost, (float*)z, (float*)m_out, (float*)minv, (float*)b_out);
#undef COV3D_MLF
  COV3D_LAUNCH_CHECK();
  return 0;
}

extern "C" int cov3d_ml_grad_images(const void* c, const void* ctf, const void* yhat,
                                    const void* z, const void* minv, const int* pix_map,
                                    int64_t npix, void* g, int64_t B, int64_t r, int64_t L,
                                    double noise_var, double grad_scale, void* stream) {
  COV3D_CHECK_ARG(r >= 1 && r <= 31, "rank r=%lld must be in [1, 31]", (long long)r);
  COV3D_CHECK_ARG(B >= 0 && L >= 1, "bad batch/size");
  COV3D_CHECK_ARG(noise_var > 0, "noise_var must be positive");
  if (B == 0) return 0;
  COV3D_CHECK_ARG(c && yhat && z && minv && g, "NULL pointer");
  COV3D_CHECK_ARG(B < 65536, "batch too large for one launch (B=%lld)", (long long)B);
  COV3D_CHECK_ARG(pix_map ? (npix >= 1 && npix <= L * L) : (npix == 0 || npix == L * L),
                  "npix=%lld invalid for L=%lld", (long long)npix, (long long)L);
  hipStream_t st = cov3d::as_stream(stream);
  const int64_t np = pix_map ? npix : L * L;
  const float scale = (float)(grad_scale / ((double)(L * L) * noise_var));
  dim3 grid((unsigned)cov3d::ceil_div(np, cov3d::kMlBlock), (unsigned)B);
#define COV3D_GRAD(RM)                                                                            \
  hipLaunchKernelGGL(cov3d::k_ml_grad<RM>, grid, dim3(cov3d::kMlBlock), 0, st, (const float2*)c, \
                     (const float*)ctf, (const float2*)yhat, (const float*)z, (const float*)minv, \
                     pix_map, np, (float2*)g, B, (int)r, (int)L, scale)
  if (r <= 4)
    COV3D_GRAD(4);
  else if (r <= 8)
    COV3D_GRAD(8);
  else if (r <= 12)
    COV3D_GRAD(12);
  else if (r <= 16)
    COV3D_GRAD(16);
  else if (r <= 20)
    COV3D_GRAD(20);
  else if (r <= 24)
    COV3D_GRAD(24);
  else
    COV3D_GRAD(32);
#undef COV3D_GRAD
  COV3D_LAUNCH_CHECK();
  return 0;
}

extern "C" int cov3d_ml_grad_sorted(const void* c, const void* ctf, const void* yhat, const void* z,
                                    const void* minv, const int* pix_map, int64_t npix, void* g_sorted,
                                    const int* spos, int64_t ldc, int64_t B, int64_t r, int64_t L,
                                    double noise_var, double grad_scale, void* stream) {
  COV3D_CHECK_ARG(r >= 1 && r <= 31, "rank r=%lld must be in [1, 31]", (long long)r);
  COV3D_CHECK_ARG(B >= 0 && L >= 1, "bad batch/size");
  COV3D_CHECK_ARG(noise_var > 0, "noise_var must be positive");
  COV3D_CHECK_ARG(ldc >= r && ldc <= cov3d::kGradLdc && ldc % 2 == 0,
                  "row stride %lld must be even, >= r and <= %d", (long long)ldc, cov3d::kGradLdc);
  if (B == 0) return 0;
  COV3D_CHECK_ARG(c && yhat && z && minv && g_sorted && spos, "NULL pointer");
  COV3D_CHECK_ARG(B < 65536, "batch too large for one launch (B=%lld)", (long long)B);
  COV3D_CHECK_ARG(pix_map ? (npix >= 1 && npix <= L * L) : (npix == 0 || npix == L * L),
                  "npix=%lld invalid for L=%lld", (long long)npix, (long long)L);
  hipStream_t st = cov3d::as_stream(stream);
  const int64_t np = pix_map ? npix : L * L;
  const float scale = (float)(grad_scale / ((double)(L * L) * noise_var));
  dim3 grid((unsigned)cov3d::ceil_div(np, cov3d::kMlBlock), (unsigned)B);
#define COV3D_GRADS(RM)                                                                                   \
  hipLaunchKernelGGL(cov3d::k_ml_grad_sorted<RM>, grid, dim3(cov3d::kMlBlock), 0, st, (const float2*)c, \
                     (const float*)ctf, (const float2*)yhat, (const float*)z, (const float*)minv, pix_map, \
                     np, (float2*)g_sorted, spos, (int)ldc, B, (int)r, (int)L, scale)
  if (r <= 4)
    COV3D_GRADS(4);
  else if (r <= 8)
    COV3D_GRADS(8);
  else if (r <= 12)
    COV3D_GRADS(12);
  else if (r <= 16)
    COV3D_GRADS(16);
  else if (r <= 20)
    COV3D_GRADS(20);
  else
    COV3D_GRADS(24);
#undef COV3D_GRADS
  COV3D_LAUNCH_CHECK();
  return 0;
}

extern "C" int cov3d_ml_grad_images_complex(const void* c, const void* ctf, const void* yhat, const void* z1,
                                            const void* minv1, const void* z2, const void* minv2,
                                            const int* pix_map, int64_t npix, void* g, int64_t B, int64_t r,
                                            int64_t L, double noise_var, double grad_scale, void* stream) {
  COV3D_CHECK_ARG(r >= 1 && r <= 31, "rank r=%lld must be in [1, 31]", (long long)r);
  COV3D_CHECK_ARG(B >= 0 && L >= 1, "bad batch/size");
  COV3D_CHECK_ARG(noise_var > 0, "noise_var must be positive");
  if (B == 0) return 0;
  COV3D_CHECK_ARG(c && yhat && z1 && minv1 && z2 && minv2 && g, "NULL pointer");
  COV3D_CHECK_ARG(g != c, "g must not alias c");
  COV3D_CHECK_ARG(B < 65536, "batch too large for one launch (B=%lld)", (long long)B);
  COV3D_CHECK_ARG(pix_map ? (npix >= 1 && npix <= L * L) : (npix == 0 || npix == L * L),
                  "npix=%lld invalid for L=%lld", (long long)npix, (long long)L);
  hipStream_t st = cov3d::as_stream(stream);
  const int64_t np = pix_map ? npix : L * L;
  const float scale = (float)(grad_scale / ((double)(L * L) * noise_var));
  dim3 grid((unsigned)cov3d::ceil_div(np, cov3d::kMlBlock), (unsigned)B);
#define COV3D_GRADC(RM)                                                                                 \
  hipLaunchKernelGGL(cov3d::k_ml_grad_cplx<RM>, grid, dim3(cov3d::kMlBlock), 0, st, (const float2*)c,   \
                     (const float*)ctf, (const float2*)yhat, (const float*)z1, (const float*)minv1,     \
                     (const float*)z2, (const float*)minv2, pix_map, np, (float2*)g, B, (int)r, (int)L, \
                     scale)
  if (r <= 4)
    COV3D_GRADC(4);
  else if (r <= 8)
    COV3D_GRADC(8);
  else if (r <= 12)
    COV3D_GRADC(12);
  else if (r <= 16)
    COV3D_GRADC(16);
  else if (r <= 20)
    COV3D_GRADC(20);
  else if (r <= 24)
    COV3D_GRADC(24);
  else
    COV3D_GRADC(32);
#undef COV3D_GRADC
  COV3D_LAUNCH_CHECK();
  return 0;
}

extern "C" int cov3d_herm_solve_c128(const void* M, const void* b, void* minv, void* z, void* logdet, int64_t n,
                                     int64_t r, void* stream) {
  COV3D_CHECK_ARG(n >= 0 && r >= 1 && r <= cov3d::kHermMax, "need n >= 0 and 1 <= r <= %d", cov3d::kHermMax);
  COV3D_CHECK_ARG(n < 2147483647, "too many matrices");
  if (n == 0) return 0;
  COV3D_CHECK_ARG(M && b && minv && z && logdet, "NULL pointer");
  hipStream_t st = cov3d::as_stream(stream);
  hipLaunchKernelGGL(cov3d::k_herm_solve, dim3((unsigned)n), dim3(cov3d::kHermThreads), 0, st, (const double2*)M,
                     (const double2*)b, (double2*)minv, (double2*)z, (double*)logdet, (int)r);
  COV3D_LAUNCH_CHECK();
  return 0;
}
