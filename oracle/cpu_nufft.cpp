// oracle/cpu_nufft.cpp -- TEST / BASELINE INFRASTRUCTURE ONLY (never the product).
//
// FP32 exponential-of-semicircle (ES) spreading and interpolation on the CPU with OpenMP, the
// spreading/interpolation half of the CPU port of the type-2 / type-1 NUFFT that bench.py times
// as its CPU baseline (the fine-grid FFTs and the deconvolution run in numpy/scipy, see
// cpu_nufft.py).  Same method as the GPU path and as the NUFFT the reference calls
// (cov3d/nufft_plan.py:263-392 through finufft): kernel phi(z) = exp(beta (sqrt(1 - (2z/w)^2) - 1)),
// w = ceil(log10(1/eps)) + 1, upsampling 2.
//
// Layouts: the fine grid is (nf, nf, nf, r) complex64 with the r transforms of a cell adjacent, and
// point values are (M, r) complex64, so every stencil tap is one contiguous r-vector update and the
// 3w kernel weights of a point serve all r transforms.  Coordinates: cpu_es_setpts converts the
// FP32 radians to fine-grid units in double and keeps, per point and axis, the first stencil cell
// (int32) and the offset of the point from it (float), so the weights carry no FP32 rounding of
// coordinates up to nf.  It also sorts the points by their first axis-0 stencil plane and, within
// a plane, by 8 x 8 blocks of their first axis-1 / axis-2 cells.
//
// Interpolation runs in parallel over points in that order, so consecutive points read
// neighbouring cells.  Spreading runs in parallel over slabs of axis-0 planes: a slab visits the
// points whose stencil meets it and updates only its own planes, so no two threads write the same
// cell and the sums need no atomics; a point whose stencil spans k slabs has its 3w weights
// evaluated k times (3w exps against w^3 r complex multiply-adds of the taps).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

inline void es_weights(float off, int w, float beta, float* wt) {
  // taps j = 0..w-1 sit at offsets (off + j) - w/2 ... in cells from the point: z = j - u, with
  // u = the point's position relative to its first stencil cell
  const float c = 4.0f / (float)(w * w);
  for (int j = 0; j < w; ++j) {
    const float z = (float)j - off;
    const float s = 1.0f - c * z * z;
    wt[j] = s > 0.f ? std::exp(beta * (std::sqrt(s) - 1.0f)) : 0.f;
  }
}

inline int wrapi(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}

}  // namespace

extern "C" {

// pts: (3, M) FP32 radians.  Outputs: base (3, M) int32 first stencil cell, off (3, M) float offset
// of the point from it (in cells), order (M) point indices bucketed by base[0] (stable), starts
// (nf + 1) starts of the planes in `order`.
void cpu_es_setpts(int nf, int w, int64_t M, const float* pts, int32_t* base, float* off, int64_t* order,
                   int64_t* starts) {
  const double scale = (double)nf / (2.0 * M_PI);
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < M; ++j) {
    for (int a = 0; a < 3; ++a) {
      double X = (double)pts[a * M + j] * scale;
      X = std::fmod(X, (double)nf);
      if (X < 0) X += nf;
      const double b = std::ceil(X - 0.5 * w);
      base[a * M + j] = wrapi((int)b, nf);
      off[a * M + j] = (float)(X - b);
    }
  }
  // counting sort by (first axis-0 cell, axis-1 block of 8, axis-2 block of 8): points of one plane
  // stay together (the spreading slabs) and neighbours within a plane touch neighbouring cells
  const int nb = (nf + 7) / 8;
  const int64_t nkey = (int64_t)nf * nb * nb;
  std::vector<int64_t> cnt(nkey + 1, 0);
  std::vector<int64_t> key(M);
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < M; ++j) key[j] = ((int64_t)base[j] * nb + base[M + j] / 8) * nb + base[2 * M + j] / 8;
  for (int64_t j = 0; j < M; ++j) cnt[key[j] + 1]++;
  for (int64_t k = 0; k < nkey; ++k) cnt[k + 1] += cnt[k];
  for (int p = 0; p <= nf; ++p) starts[p] = cnt[(int64_t)p * nb * nb];
  for (int64_t j = 0; j < M; ++j) order[cnt[key[j]]++] = j;
}

// Type 2 interpolation: out[j, t] = sum over the w^3 stencil of phi phi phi * grid[cell, t].
void cpu_es_interp_f32(int nf, int w, float beta, int64_t M, const int32_t* base, const float* off,
                       const int64_t* order, const float* grid, int r, float* out) {
  const int64_t nf2 = (int64_t)nf * nf;
#pragma omp parallel
  {
    std::vector<float> acc(2 * r);
#pragma omp for schedule(dynamic, 4096)
    for (int64_t q = 0; q < M; ++q) {
      const int64_t j = order[q];
      float w0[16], w1[16], w2[16];
      es_weights(off[j], w, beta, w0);
      es_weights(off[M + j], w, beta, w1);
      es_weights(off[2 * M + j], w, beta, w2);
      const int b0 = base[j], b1 = base[M + j], b2 = base[2 * M + j];
      std::fill(acc.begin(), acc.end(), 0.f);
      float* a = acc.data();
      for (int i0 = 0; i0 < w; ++i0) {
        const int64_t p0 = (int64_t)wrapi(b0 + i0, nf) * nf2;
        for (int i1 = 0; i1 < w; ++i1) {
          const int64_t p1 = p0 + (int64_t)wrapi(b1 + i1, nf) * nf;
          const float w01 = w0[i0] * w1[i1];
          for (int i2 = 0; i2 < w; ++i2) {
            const float wt = w01 * w2[i2];
            const float* g = grid + 2 * (p1 + wrapi(b2 + i2, nf)) * r;
#pragma omp simd
            for (int t = 0; t < 2 * r; ++t) a[t] += wt * g[t];
          }
        }
      }
      std::memcpy(out + 2 * j * r, a, sizeof(float) * 2 * r);
    }
  }
}

// Type 1 spreading: grid[cell, t] = sum over points whose stencil covers the cell of phi phi phi *
// c[j, t].  The grid is cleared here.  Slabs of `slab` axis-0 planes are the parallel units.
void cpu_es_spread_f32(int nf, int w, float beta, int64_t M, const int32_t* base, const float* off,
                       const int64_t* order, const int64_t* starts, const float* c, int r, float* grid,
                       int slab) {
  const int64_t nf2 = (int64_t)nf * nf;
  const int nslab = (nf + slab - 1) / slab;
#pragma omp parallel for schedule(dynamic, 1)
  for (int s = 0; s < nslab; ++s) {
    const int a0 = s * slab, a1 = std::min(nf, a0 + slab);
    std::memset(grid + 2 * (int64_t)a0 * nf2 * r, 0, sizeof(float) * 2 * (int64_t)(a1 - a0) * nf2 * r);
    // points whose first axis-0 cell lies in [a0 - w + 1, a1 - 1] (mod nf) reach this slab
    for (int pb = a0 - w + 1; pb < a1; ++pb) {
      const int p = wrapi(pb, nf);
      for (int64_t q = starts[p]; q < starts[p + 1]; ++q) {
        const int64_t j = order[q];
        float w0[16], w1[16], w2[16];
        es_weights(off[j], w, beta, w0);
        es_weights(off[M + j], w, beta, w1);
        es_weights(off[2 * M + j], w, beta, w2);
        const int b1 = base[M + j], b2 = base[2 * M + j];
        const float* v = c + 2 * j * r;
        for (int i0 = 0; i0 < w; ++i0) {
          const int plane = wrapi(pb + i0, nf);
          if (plane < a0 || plane >= a1) continue;
          const int64_t p0 = (int64_t)plane * nf2;
          for (int i1 = 0; i1 < w; ++i1) {
            const int64_t p1 = p0 + (int64_t)wrapi(b1 + i1, nf) * nf;
            const float w01 = w0[i0] * w1[i1];
            for (int i2 = 0; i2 < w; ++i2) {
              const float wt = w01 * w2[i2];
              float* g = grid + 2 * (p1 + wrapi(b2 + i2, nf)) * r;
#pragma omp simd
              for (int t = 0; t < 2 * r; ++t) g[t] += wt * v[t];
            }
          }
        }
      }
    }
  }
}

int cpu_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
