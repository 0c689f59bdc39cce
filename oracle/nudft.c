/*
 * oracle/nudft.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * FP64 direct non-uniform DFTs in 3-D with the conventions of the
 * reference's NUFFT plans (cov3d/nufft_plan.py:263-392, finufft defaults):
 *   type 2: c[t, j] = sum_n f[t, n] exp(iflag * i * (n0 x_j + n1 y_j + n2 z_j)),
 *   type 1: f[t, n] = sum_j c[t, j] exp(iflag * i * (n0 x_j + n1 y_j + n2 z_j)),
 *   n_k in [-floor(N_k/2), ceil(N_k/2) - 1] stored at index n_k + floor(N_k/2)
 *   (C order, axis 0 slowest); complex data interleaved (re, im).
 * Cost O(M * N0 * N1 * N2 * ntrans): tiny sizes only.
 *
 * Also a CPU exponential-of-semicircle (ES) spreader / interpolator used by
 * oracle/cpu_nufft.py as the CPU port of the NUFFT for the baseline timing.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

static void axis_phases(int N, double x, int iflag, cplx* e) {
  /* e[i] = exp(iflag * i * n * x), n = i - N/2, by recurrence from a direct start */
  int h = N / 2;
  cplx step = cexp(I * (double)iflag * x);
  cplx v = cexp(I * (double)iflag * (double)(-h) * x);
  for (int i = 0; i < N; ++i) {
    e[i] = v;
    v *= step;
    if ((i & 15) == 15) v = cexp(I * (double)iflag * (double)(i + 1 - h) * x);  /* limit drift */
  }
}

void oracle_nudft3d_type2(int N0, int N1, int N2, int64_t M, const double* x, const double* y,
                          const double* z, const double* f, int ntrans, int iflag, double* c) {
  const int64_t nm = (int64_t)N0 * N1 * N2;
  const cplx* F = (const cplx*)f;
  cplx* C = (cplx*)c;
#pragma omp parallel
  {
    cplx* e0 = malloc(sizeof(cplx) * N0);
    cplx* e1 = malloc(sizeof(cplx) * N1);
    cplx* e2 = malloc(sizeof(cplx) * N2);
#pragma omp for schedule(dynamic, 16)
    for (int64_t j = 0; j < M; ++j) {
      axis_phases(N0, x[j], iflag, e0);
      axis_phases(N1, y[j], iflag, e1);
      axis_phases(N2, z[j], iflag, e2);
      for (int t = 0; t < ntrans; ++t) {
        const cplx* ft = F + (int64_t)t * nm;
        cplx s0 = 0;
        for (int i0 = 0; i0 < N0; ++i0) {
          cplx s1 = 0;
          for (int i1 = 0; i1 < N1; ++i1) {
            const cplx* row = ft + ((int64_t)i0 * N1 + i1) * N2;
            cplx s2 = 0;
            for (int i2 = 0; i2 < N2; ++i2) s2 += e2[i2] * row[i2];
            s1 += e1[i1] * s2;
          }
          s0 += e0[i0] * s1;
        }
        C[(int64_t)t * M + j] = s0;
      }
    }
    free(e0);
    free(e1);
    free(e2);
  }
}

void oracle_nudft3d_type1(int N0, int N1, int N2, int64_t M, const double* x, const double* y,
                          const double* z, const double* c, int ntrans, int iflag, double* f) {
  const int64_t nm = (int64_t)N0 * N1 * N2;
  const cplx* C = (const cplx*)c;
  cplx* F = (cplx*)f;
  memset(F, 0, sizeof(cplx) * nm * ntrans);
#pragma omp parallel
  {
    cplx* e0 = malloc(sizeof(cplx) * N0);
    cplx* e1 = malloc(sizeof(cplx) * N1);
    cplx* e2 = malloc(sizeof(cplx) * N2);
    cplx* acc = calloc((size_t)nm * ntrans, sizeof(cplx));
#pragma omp for schedule(dynamic, 16)
    for (int64_t j = 0; j < M; ++j) {
      axis_phases(N0, x[j], iflag, e0);
      axis_phases(N1, y[j], iflag, e1);
      axis_phases(N2, z[j], iflag, e2);
      for (int t = 0; t < ntrans; ++t) {
        cplx cj = C[(int64_t)t * M + j];
        cplx* ft = acc + (int64_t)t * nm;
        for (int i0 = 0; i0 < N0; ++i0) {
          cplx a0 = cj * e0[i0];
          for (int i1 = 0; i1 < N1; ++i1) {
            cplx a1 = a0 * e1[i1];
            cplx* row = ft + ((int64_t)i0 * N1 + i1) * N2;
            for (int i2 = 0; i2 < N2; ++i2) row[i2] += a1 * e2[i2];
          }
        }
      }
    }
#pragma omp critical
    for (int64_t k = 0; k < nm * ntrans; ++k) F[k] += acc[k];
    free(acc);
    free(e0);
    free(e1);
    free(e2);
  }
}

/* ---------------------------------------------------------------------------
 * CPU ES spreading / interpolation on a periodic (nf0, nf1, nf2) complex grid.
 * X, Y, Z are fine-grid coordinates in [0, nf) (x * nf / (2 pi) wrapped).
 * ------------------------------------------------------------------------- */
static inline double es(double t, int w, double beta) {
  double c = 4.0 / ((double)w * w);
  double s = 1.0 - c * t * t;
  return s > 0 ? exp(beta * (sqrt(s) - 1.0)) : 0.0;
}

static inline void weights(double X, int w, double beta, int* i0, double* wt) {
  int j0 = (int)ceil(X - 0.5 * w);
  *i0 = j0;
  for (int a = 0; a < w; ++a) wt[a] = es((double)(j0 + a) - X, w, beta);
}

static inline int wrapi(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}

void oracle_es_interp(int nf0, int nf1, int nf2, int64_t M, const double* X, const double* Y,
                      const double* Z, const double* grid, int ntrans, int w, double beta,
                      double* c) {
  const int64_t ng = (int64_t)nf0 * nf1 * nf2;
  const cplx* G = (const cplx*)grid;
  cplx* C = (cplx*)c;
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < M; ++j) {
    double w0[16], w1[16], w2[16];
    int i0, i1, i2;
    weights(X[j], w, beta, &i0, w0);
    weights(Y[j], w, beta, &i1, w1);
    weights(Z[j], w, beta, &i2, w2);
    for (int t = 0; t < ntrans; ++t) {
      const cplx* g = G + (int64_t)t * ng;
      cplx s = 0;
      for (int a = 0; a < w; ++a) {
        int m0 = wrapi(i0 + a, nf0);
        for (int b = 0; b < w; ++b) {
          int m1 = wrapi(i1 + b, nf1);
          double wab = w0[a] * w1[b];
          const cplx* row = g + ((int64_t)m0 * nf1 + m1) * nf2;
          for (int cc = 0; cc < w; ++cc) s += wab * w2[cc] * row[wrapi(i2 + cc, nf2)];
        }
      }
      C[(int64_t)t * M + j] = s;
    }
  }
}

void oracle_es_spread(int nf0, int nf1, int nf2, int64_t M, const double* X, const double* Y,
                      const double* Z, const double* c, int ntrans, int w, double beta,
                      double* grid) {
  const int64_t ng = (int64_t)nf0 * nf1 * nf2;
  const cplx* C = (const cplx*)c;
  cplx* G = (cplx*)grid;
  memset(G, 0, sizeof(cplx) * ng * ntrans);
  /* one transform per thread: no write conflicts */
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < ntrans; ++t) {
    cplx* g = G + (int64_t)t * ng;
    for (int64_t j = 0; j < M; ++j) {
      double w0[16], w1[16], w2[16];
      int i0, i1, i2;
      weights(X[j], w, beta, &i0, w0);
      weights(Y[j], w, beta, &i1, w1);
      weights(Z[j], w, beta, &i2, w2);
      cplx cj = C[(int64_t)t * M + j];
      for (int a = 0; a < w; ++a) {
        int m0 = wrapi(i0 + a, nf0);
        for (int b = 0; b < w; ++b) {
          int m1 = wrapi(i1 + b, nf1);
          cplx v = cj * (w0[a] * w1[b]);
          cplx* row = g + ((int64_t)m0 * nf1 + m1) * nf2;
          for (int cc = 0; cc < w; ++cc) row[wrapi(i2 + cc, nf2)] += v * w2[cc];
        }
      }
    }
  }
}
