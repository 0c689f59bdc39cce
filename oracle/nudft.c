/*
 * oracle/nudft.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * FP64 direct non-uniform DFTs in 3-D with the conventions of the
 * reference's NUFFT plans (cov3d/nufft_plan.py:263-392, finufft defaults):
 *   type 2: c[t, j] = sum_n f[t, n] exp(iflag * i * (n0 x_j + n1 y_j + n2 z_j)),
 *   type 1: f[t, n] = sum_j c[t, j] exp(iflag * i * (n0 x_j + n1 y_j + n2 z_j)),
 *   n_k in [-floor(N_k/2), ceil(N_k/2) - 1] stored at index n_k + floor(N_k/2)
 *   (C order, axis 0 slowest); complex data interleaved (re, im).
 * Cost O(M * N0 * N1 * N2 * ntrans): a few thousand points at N = 128 take seconds on a
 * few cores; type 1 can also be evaluated at a chosen subset of output modes only.
 *
 * Also a CPU exponential-of-semicircle (ES) spreader / interpolator used by
 * oracle/cpu_nufft.py as the CPU port of the NUFFT for the baseline timing.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

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

/* Points are taken in blocks of PB per thread: every volume row (one (t, i0, i1) line of N2
 * modes) is read once per block and reused for all PB points, so large volumes stay cheap in
 * memory traffic; the arithmetic is the plain triple sum. */
#define PB 32
void oracle_nudft3d_type2(int N0, int N1, int N2, int64_t M, const double* x, const double* y,
                          const double* z, const double* f, int ntrans, int iflag, double* c) {
  const int64_t nm = (int64_t)N0 * N1 * N2;
  const cplx* F = (const cplx*)f;
  cplx* C = (cplx*)c;
  const int64_t nblk = (M + PB - 1) / PB;
#pragma omp parallel
  {
    cplx* e0 = malloc(sizeof(cplx) * N0 * PB);
    cplx* e1 = malloc(sizeof(cplx) * N1 * PB);
    double* e2r = malloc(sizeof(double) * N2 * PB);
    double* e2i = malloc(sizeof(double) * N2 * PB);
    cplx* tmp = malloc(sizeof(cplx) * N2);
#pragma omp for schedule(dynamic, 1) collapse(2)
    for (int64_t blk = 0; blk < nblk; ++blk) {
      for (int t = 0; t < ntrans; ++t) {
        const int64_t j0 = blk * PB;
        const int np = (int)((M - j0) < PB ? (M - j0) : PB);
        for (int p = 0; p < np; ++p) {
          axis_phases(N0, x[j0 + p], iflag, e0 + (int64_t)p * N0);
          axis_phases(N1, y[j0 + p], iflag, e1 + (int64_t)p * N1);
          axis_phases(N2, z[j0 + p], iflag, tmp);
          for (int i2 = 0; i2 < N2; ++i2) {
            e2r[p * N2 + i2] = creal(tmp[i2]);
            e2i[p * N2 + i2] = cimag(tmp[i2]);
          }
        }
        const cplx* ft = F + (int64_t)t * nm;
        cplx s0[PB], s1[PB];
        for (int p = 0; p < np; ++p) s0[p] = 0;
        for (int i0 = 0; i0 < N0; ++i0) {
          for (int p = 0; p < np; ++p) s1[p] = 0;
          for (int i1 = 0; i1 < N1; ++i1) {
            const double* row = (const double*)(ft + ((int64_t)i0 * N1 + i1) * N2);
            for (int p = 0; p < np; ++p) {
              const double* er = e2r + p * N2;
              const double* ei = e2i + p * N2;
              double ar[4] = {0, 0, 0, 0}, ai[4] = {0, 0, 0, 0};
              int i2 = 0;
              for (; i2 + 4 <= N2; i2 += 4) {
                for (int u = 0; u < 4; ++u) {
                  const double vr = row[2 * (i2 + u)], vi = row[2 * (i2 + u) + 1];
                  ar[u] += er[i2 + u] * vr - ei[i2 + u] * vi;
                  ai[u] += er[i2 + u] * vi + ei[i2 + u] * vr;
                }
              }
              for (; i2 < N2; ++i2) {
                const double vr = row[2 * i2], vi = row[2 * i2 + 1];
                ar[0] += er[i2] * vr - ei[i2] * vi;
                ai[0] += er[i2] * vi + ei[i2] * vr;
              }
              const double sr = (ar[0] + ar[1]) + (ar[2] + ar[3]);
              const double si = (ai[0] + ai[1]) + (ai[2] + ai[3]);
              s1[p] += e1[(int64_t)p * N1 + i1] * (sr + I * si);
            }
          }
          for (int p = 0; p < np; ++p) s0[p] += e0[(int64_t)p * N0 + i0] * s1[p];
        }
        for (int p = 0; p < np; ++p) C[(int64_t)t * M + j0 + p] = s0[p];
      }
    }
    free(e0);
    free(e1);
    free(e2r);
    free(e2i);
    free(tmp);
  }
}

/* Type 1 at a subset of output modes: out[t, q] = sum_j c[t, j] exp(iflag * i * n_q . p_j) for
 * the flat grid indices modes[q] (C order over (N0, N1, N2), n = index - N/2 per axis).  Each
 * thread sums a range of points into its own (ntrans, nq) accumulator; the partial sums are
 * added in thread order, so the result does not depend on the schedule's timing. */
void oracle_nudft3d_type1_modes(int N0, int N1, int N2, int64_t M, const double* x, const double* y,
                                const double* z, const double* c, int ntrans, int iflag, int nq,
                                const int64_t* modes, double* out) {
  const cplx* C = (const cplx*)c;
  cplx* O = (cplx*)out;
  memset(O, 0, sizeof(cplx) * (size_t)nq * ntrans);
  double* n0 = malloc(sizeof(double) * nq);
  double* n1 = malloc(sizeof(double) * nq);
  double* n2 = malloc(sizeof(double) * nq);
  for (int q = 0; q < nq; ++q) {
    int64_t m = modes[q];
    n2[q] = (double)(m % N2) - (double)(N2 / 2);
    n1[q] = (double)((m / N2) % N1) - (double)(N1 / 2);
    n0[q] = (double)(m / ((int64_t)N1 * N2)) - (double)(N0 / 2);
  }
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  cplx* part = calloc((size_t)nth * nq * ntrans, sizeof(cplx));
#pragma omp parallel
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    cplx* acc = part + (size_t)tid * nq * ntrans;
    cplx* e = malloc(sizeof(cplx) * nq);
#pragma omp for schedule(static)
    for (int64_t j = 0; j < M; ++j) {
      for (int q = 0; q < nq; ++q) e[q] = cexp(I * (double)iflag * (n0[q] * x[j] + n1[q] * y[j] + n2[q] * z[j]));
      for (int t = 0; t < ntrans; ++t) {
        const cplx cj = C[(int64_t)t * M + j];
        cplx* a = acc + (size_t)t * nq;
        for (int q = 0; q < nq; ++q) a[q] += cj * e[q];
      }
    }
    free(e);
  }
  for (int th = 0; th < nth; ++th)
    for (size_t k = 0; k < (size_t)nq * ntrans; ++k) O[k] += part[(size_t)th * nq * ntrans + k];
  free(part);
  free(n0);
  free(n1);
  free(n2);
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
