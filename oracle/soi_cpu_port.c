/*
 * soi_cpu_port.c -- CPU performance port of the reference's timed region (TEST INFRASTRUCTURE:
 * the CPU baseline that bench.py times beside the GPU; never linked by the product).
 *
 * What it times is what test.c:389-404 times: compute_soi on an input already in memory, the
 * descriptor (windows, W_inv, FFT tables) built beforehand.  The P ranks run in one address
 * space with OpenMP threads in place of MPI ranks, so the all-to-all (pfft.c:544-583) is the
 * segment-major layout of alpha_tilde, as on the GPU.  The arithmetic is the reference's:
 *   Step 1 (parallel_filter_subsampling.cpp:134-311): complex-tap B-tap convolution per lane
 *     i with the reference's window values w[theta][b][i] (orc_w_table), 8 flops per tap and
 *     output, vectorised over the S lanes of a row, then the S-point DFT of each row;
 *   Step 2 (pfft.c:642): the M_hat-point forward FFT of each segment, here as a four-step
 *     (M_hat = n1 * n2) mixed-radix Stockham FFT with precomputed twiddle tables;
 *   demodulation X[s*M + m] = W_inv[m] * y[s][m] (pfft.c:649-693), fused into the last pass.
 * Differences from the reference's CPU code: no MKL (the FFTs are this file's), no MPI.  The
 * output matches orc_compute_soi to rounding (tests/test_oracle.py).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <immintrin.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  long N;
  int P, k, n_mu, d_mu, B;
  double tau, sigma;
} orc_params; /* the oracle's parameter block (soi_oracle.c) */

void orc_w_table(const orc_params *p, double *w);
void orc_w_inv_table(const orc_params *p, double *W_inv);

static const long double PI_L = 3.141592653589793238462643383279502884197169399375L;

/* e^{-2 pi i num/den} */
static void tw_exact(long num, long den, double *c, double *s)
{
  long r = num % den;
  if (r < 0) r += den;
  long double a = -2 * PI_L * (long double)r / (long double)den;
  *c = (double)cosl(a);
  *s = (double)sinl(a);
}

/* ---- Stockham autosort FFT of length L over a batch of NB interleaved sequences ----------
 * Data layout [L][NB] (the batch index fastest), so every butterfly's inner loop runs over
 * NB * stride contiguous complex values.  Stage with radix r on current length n and stride s
 * (decimation in frequency): for p < m = n/r, q < s, the r inputs x[q + s(p + j m)] give
 * y[q + s(r p + k)] = w_n^{k p} * sum_j x_j w_r^{jk}. */
#define PORT_MAXST 24
typedef struct {
  int L, nst;
  int r[PORT_MAXST];
  double *tw[PORT_MAXST]; /* per stage: (r - 1) * m twiddles w_n^{k p}, [p][k-1] */
} port_fft;

static int port_fft_init(port_fft *f, int L)
{
  memset(f, 0, sizeof(*f));
  f->L = L;
  int n = L;
  while (n > 1) {
    int r = n % 4 == 0 ? 4 : n % 2 == 0 ? 2 : n % 5 == 0 ? 5 : n % 3 == 0 ? 3 : 0;
    if (!r || f->nst == PORT_MAXST) return -1;
    int m = n / r;
    double *t = (double *)malloc(sizeof(double) * 2 * (size_t)(r - 1) * (m > 0 ? m : 1));
    for (int p = 0; p < m; ++p)
      for (int k = 1; k < r; ++k) tw_exact((long)k * p, n, &t[2 * (p * (r - 1) + k - 1)],
                                           &t[2 * (p * (r - 1) + k - 1) + 1]);
    f->r[f->nst] = r;
    f->tw[f->nst] = t;
    f->nst++;
    n = m;
  }
  return 0;
}

static void port_fft_free(port_fft *f)
{
  for (int s = 0; s < f->nst; ++s) free(f->tw[s]);
}

/* radix-r butterfly over complex vectors of length V (contiguous, interleaved) */
static inline void bfly(int r, double *restrict const *xin, double *restrict const *yout,
                        const double *tw, long V)
{
  const double c3 = -0.5, s3 = -0.86602540378443864676;
  const double c51 = 0.30901699437494742410, c52 = -0.80901699437494742410;
  const double s51 = -0.95105651629515357212, s52 = -0.58778525229247312917;
  if (r == 2) {
    const double wr = tw[0], wi = tw[1];
#pragma omp simd
    for (long v = 0; v < V; ++v) {
      double ar = xin[0][2 * v], ai = xin[0][2 * v + 1], br = xin[1][2 * v], bi = xin[1][2 * v + 1];
      yout[0][2 * v] = ar + br;
      yout[0][2 * v + 1] = ai + bi;
      double dr = ar - br, di = ai - bi;
      yout[1][2 * v] = dr * wr - di * wi;
      yout[1][2 * v + 1] = dr * wi + di * wr;
    }
  } else if (r == 4) {
    const double w1r = tw[0], w1i = tw[1], w2r = tw[2], w2i = tw[3], w3r = tw[4], w3i = tw[5];
#pragma omp simd
    for (long v = 0; v < V; ++v) {
      double ar = xin[0][2 * v], ai = xin[0][2 * v + 1], br = xin[1][2 * v], bi = xin[1][2 * v + 1];
      double cr = xin[2][2 * v], ci = xin[2][2 * v + 1], dr = xin[3][2 * v], di = xin[3][2 * v + 1];
      double t0r = ar + cr, t0i = ai + ci, t1r = ar - cr, t1i = ai - ci;
      double t2r = br + dr, t2i = bi + di, t3r = bi - di, t3i = dr - br; /* -i (b - d) */
      yout[0][2 * v] = t0r + t2r;
      yout[0][2 * v + 1] = t0i + t2i;
      double y1r = t1r + t3r, y1i = t1i + t3i, y2r = t0r - t2r, y2i = t0i - t2i;
      double y3r = t1r - t3r, y3i = t1i - t3i;
      yout[1][2 * v] = y1r * w1r - y1i * w1i;
      yout[1][2 * v + 1] = y1r * w1i + y1i * w1r;
      yout[2][2 * v] = y2r * w2r - y2i * w2i;
      yout[2][2 * v + 1] = y2r * w2i + y2i * w2r;
      yout[3][2 * v] = y3r * w3r - y3i * w3i;
      yout[3][2 * v + 1] = y3r * w3i + y3i * w3r;
    }
  } else if (r == 3) {
    const double w1r = tw[0], w1i = tw[1], w2r = tw[2], w2i = tw[3];
#pragma omp simd
    for (long v = 0; v < V; ++v) {
      double ar = xin[0][2 * v], ai = xin[0][2 * v + 1], br = xin[1][2 * v], bi = xin[1][2 * v + 1];
      double cr = xin[2][2 * v], ci = xin[2][2 * v + 1];
      double sr = br + cr, si = bi + ci, dr = br - cr, di = bi - ci;
      double mr = ar + c3 * sr, mi = ai + c3 * si;
      yout[0][2 * v] = ar + sr;
      yout[0][2 * v + 1] = ai + si;
      /* X1 = m + i s3 d, X2 = m - i s3 d (s3 = -sin(2pi/3)) */
      double y1r = mr - s3 * di, y1i = mi + s3 * dr, y2r = mr + s3 * di, y2i = mi - s3 * dr;
      yout[1][2 * v] = y1r * w1r - y1i * w1i;
      yout[1][2 * v + 1] = y1r * w1i + y1i * w1r;
      yout[2][2 * v] = y2r * w2r - y2i * w2i;
      yout[2][2 * v + 1] = y2r * w2i + y2i * w2r;
    }
  } else { /* r == 5 */
    const double w1r = tw[0], w1i = tw[1], w2r = tw[2], w2i = tw[3], w3r = tw[4], w3i = tw[5];
    const double w4r = tw[6], w4i = tw[7];
#pragma omp simd
    for (long v = 0; v < V; ++v) {
      double x0r = xin[0][2 * v], x0i = xin[0][2 * v + 1];
      double b1r = xin[1][2 * v] + xin[4][2 * v], b1i = xin[1][2 * v + 1] + xin[4][2 * v + 1];
      double b2r = xin[2][2 * v] + xin[3][2 * v], b2i = xin[2][2 * v + 1] + xin[3][2 * v + 1];
      double d1r = xin[1][2 * v] - xin[4][2 * v], d1i = xin[1][2 * v + 1] - xin[4][2 * v + 1];
      double d2r = xin[2][2 * v] - xin[3][2 * v], d2i = xin[2][2 * v + 1] - xin[3][2 * v + 1];
      double p1r = x0r + c51 * b1r + c52 * b2r, p1i = x0i + c51 * b1i + c52 * b2i;
      double p2r = x0r + c52 * b1r + c51 * b2r, p2i = x0i + c52 * b1i + c51 * b2i;
      /* q = i * (s51 d1 + s52 d2) etc. with s5k = -sin(2 pi k/5) */
      double q1r = -(s51 * d1i + s52 * d2i), q1i = s51 * d1r + s52 * d2r;
      double q2r = -(s52 * d1i - s51 * d2i), q2i = s52 * d1r - s51 * d2r;
      yout[0][2 * v] = x0r + b1r + b2r;
      yout[0][2 * v + 1] = x0i + b1i + b2i;
      double y1r = p1r + q1r, y1i = p1i + q1i, y4r = p1r - q1r, y4i = p1i - q1i;
      double y2r = p2r + q2r, y2i = p2i + q2i, y3r = p2r - q2r, y3i = p2i - q2i;
      yout[1][2 * v] = y1r * w1r - y1i * w1i;
      yout[1][2 * v + 1] = y1r * w1i + y1i * w1r;
      yout[2][2 * v] = y2r * w2r - y2i * w2i;
      yout[2][2 * v + 1] = y2r * w2i + y2i * w2r;
      yout[3][2 * v] = y3r * w3r - y3i * w3i;
      yout[3][2 * v + 1] = y3r * w3i + y3i * w3r;
      yout[4][2 * v] = y4r * w4r - y4i * w4i;
      yout[4][2 * v + 1] = y4r * w4i + y4i * w4r;
    }
  }
}

/* FFT of NB sequences, layout [L][NB]; x and work hold L*NB complex; returns the buffer that
 * holds the result (x or work) */
static double *port_fft_run(const port_fft *f, double *x, double *work, long NB)
{
  static const double one[8] = {1, 0, 1, 0, 1, 0, 1, 0};
  double *src = x, *dst = work;
  long n = f->L, s = 1;
  for (int st = 0; st < f->nst; ++st) {
    const int r = f->r[st];
    const long m = n / r;
    for (long p = 0; p < m; ++p) {
      double *xin[5], *yout[5];
      for (int j = 0; j < r; ++j) {
        xin[j] = src + 2 * (s * (p + j * m)) * NB;
        yout[j] = dst + 2 * (s * (r * p + j)) * NB;
      }
      bfly(r, xin, yout, p == 0 ? one : f->tw[st] + 2 * p * (r - 1), s * NB);
    }
    double *t = src;
    src = dst;
    dst = t;
    n = m;
    s *= r;
  }
  return src;
}

/* ---- the port's plan ---------------------------------------------------------------------- */
typedef struct {
  orc_params p;
  long S, M, Mh, nJ;
  double *w;      /* [theta][b][S] complex: the reference's window values */
  double *wdr, *wdi; /* the same as duplicated real and imaginary parts (the reference's w_dup,
                        pfft.c:198-204): [theta][b][S] pairs (re, re) and (im, im) */
  double *winv;   /* [M] complex */
  int n1, n2;     /* four-step split of M_hat */
  port_fft fS, f1, f2;
  double *thi, *tlo; /* w_Mh^e = thi[e >> 10] * tlo[e & 1023] */
  double *at;     /* alpha_tilde [S][M_hat] */
} port_plan;

#define PORT_JB 16 /* row groups per filter block */
#define PORT_CB 8  /* columns per pass-1 block, rows per pass-2 block */

void *port_create(const orc_params *p)
{
  port_plan *pl = (port_plan *)calloc(1, sizeof(port_plan));
  pl->p = *p;
  pl->S = (long)p->k * p->P;
  pl->M = p->N / pl->S;
  pl->Mh = (long)p->n_mu * pl->M / p->d_mu;
  pl->nJ = pl->M / p->d_mu;
  pl->w = (double *)malloc(sizeof(double) * 2 * (size_t)p->n_mu * p->B * pl->S);
  pl->winv = (double *)malloc(sizeof(double) * 2 * (size_t)pl->M);
  orc_w_table(p, pl->w);
  orc_w_inv_table(p, pl->winv);
  {
    const size_t nw = (size_t)p->n_mu * p->B * pl->S;
    pl->wdr = (double *)aligned_alloc(64, sizeof(double) * 2 * nw);
    pl->wdi = (double *)aligned_alloc(64, sizeof(double) * 2 * nw);
    for (size_t e = 0; e < nw; ++e) {
      pl->wdr[2 * e] = pl->wdr[2 * e + 1] = pl->w[2 * e];
      pl->wdi[2 * e] = pl->wdi[2 * e + 1] = pl->w[2 * e + 1];
    }
  }
  /* n2 = the largest power of two <= sqrt(M_hat) that divides it */
  long n2 = 1;
  while (n2 * n2 * 4 <= pl->Mh && pl->Mh % (n2 * 2) == 0) n2 *= 2;
  pl->n2 = (int)n2;
  pl->n1 = (int)(pl->Mh / n2);
  if (port_fft_init(&pl->fS, (int)pl->S) || port_fft_init(&pl->f1, pl->n1) ||
      port_fft_init(&pl->f2, pl->n2)) {
    free(pl->w);
    free(pl->winv);
    free(pl);
    return NULL;
  }
  long nhi = (pl->Mh + 1023) / 1024;
  pl->thi = (double *)malloc(sizeof(double) * 2 * nhi);
  pl->tlo = (double *)malloc(sizeof(double) * 2 * 1024);
  for (long e = 0; e < nhi; ++e) tw_exact(e * 1024, pl->Mh, &pl->thi[2 * e], &pl->thi[2 * e + 1]);
  for (long e = 0; e < 1024; ++e) tw_exact(e, pl->Mh, &pl->tlo[2 * e], &pl->tlo[2 * e + 1]);
  pl->at = (double *)malloc(sizeof(double) * 2 * (size_t)pl->S * pl->Mh);
  return pl;
}

void port_destroy(void *h)
{
  port_plan *pl = (port_plan *)h;
  if (!pl) return;
  port_fft_free(&pl->fS);
  port_fft_free(&pl->f1);
  port_fft_free(&pl->f2);
  free(pl->w);
  free(pl->wdr);
  free(pl->wdi);
  free(pl->winv);
  free(pl->thi);
  free(pl->tlo);
  free(pl->at);
  free(pl);
}

/* Step 1: blocks of PORT_JB row groups.  gamma of the block's R rows goes to g[i][R'] (lane
 * major), the S-point DFT runs over the R' batch, and each segment's R' run is stored
 * contiguously into alpha_tilde[s][R]. */
static void port_filter(port_plan *pl, const double *x)
{
  const long S = pl->S, M = pl->M, Mh = pl->Mh, nJ = pl->nJ;
  const int n_mu = pl->p.n_mu, d_mu = pl->p.d_mu, B = pl->p.B;
  const long nblk = (nJ + PORT_JB - 1) / PORT_JB;
  const long RB = (long)PORT_JB * n_mu;
#pragma omp parallel
  {
    double *acc = (double *)malloc(sizeof(double) * 2 * S);
    double *g = (double *)malloc(sizeof(double) * 2 * S * RB);
    double *g2 = (double *)malloc(sizeof(double) * 2 * S * RB);
#pragma omp for schedule(static)
    for (long blk = 0; blk < nblk; ++blk) {
      const long J0 = blk * PORT_JB;
      const long nj = J0 + PORT_JB <= nJ ? PORT_JB : nJ - J0;
      for (long jj = 0; jj < nj; ++jj) {
        const long J = J0 + jj;
        for (int th = 0; th < n_mu; ++th) {
          const long rr = jj * n_mu + th;
          if (S % 8 == 0) {
            /* AVX2: two complex per vector; a = sum wr*x, c = sum wi*swap(x), then
             * gamma = addsub(a, c) = (sum wr xr - wi xi, sum wr xi + wi xr) */
            for (long i0 = 0; i0 < S; i0 += 8) {
              __m256d a0 = _mm256_setzero_pd(), a1 = a0, a2 = a0, a3 = a0;
              __m256d c0 = a0, c1 = a0, c2 = a0, c3 = a0;
              for (int b = 0; b < B; ++b) {
                long q = J * d_mu + b;
                if (q >= M) q -= M;
                const double *xr = x + 2 * (q * S + i0);
                const long wo = 2 * (((long)th * B + b) * S + i0);
                const double *wr = pl->wdr + wo, *wi = pl->wdi + wo;
                const __m256d x0 = _mm256_loadu_pd(xr), x1 = _mm256_loadu_pd(xr + 4);
                const __m256d x2 = _mm256_loadu_pd(xr + 8), x3 = _mm256_loadu_pd(xr + 12);
                a0 = _mm256_fmadd_pd(_mm256_load_pd(wr), x0, a0);
                a1 = _mm256_fmadd_pd(_mm256_load_pd(wr + 4), x1, a1);
                a2 = _mm256_fmadd_pd(_mm256_load_pd(wr + 8), x2, a2);
                a3 = _mm256_fmadd_pd(_mm256_load_pd(wr + 12), x3, a3);
                c0 = _mm256_fmadd_pd(_mm256_load_pd(wi), _mm256_permute_pd(x0, 5), c0);
                c1 = _mm256_fmadd_pd(_mm256_load_pd(wi + 4), _mm256_permute_pd(x1, 5), c1);
                c2 = _mm256_fmadd_pd(_mm256_load_pd(wi + 8), _mm256_permute_pd(x2, 5), c2);
                c3 = _mm256_fmadd_pd(_mm256_load_pd(wi + 12), _mm256_permute_pd(x3, 5), c3);
              }
              double o[16];
              _mm256_storeu_pd(o, _mm256_addsub_pd(a0, c0));
              _mm256_storeu_pd(o + 4, _mm256_addsub_pd(a1, c1));
              _mm256_storeu_pd(o + 8, _mm256_addsub_pd(a2, c2));
              _mm256_storeu_pd(o + 12, _mm256_addsub_pd(a3, c3));
              for (int e = 0; e < 8; ++e) {
                g[2 * ((i0 + e) * RB + rr)] = o[2 * e];
                g[2 * ((i0 + e) * RB + rr) + 1] = o[2 * e + 1];
              }
            }
          } else {
            memset(acc, 0, sizeof(double) * 2 * S);
            for (int b = 0; b < B; ++b) {
              long q = J * d_mu + b;
              if (q >= M) q -= M;
              const double *xr = x + 2 * q * S;
              const double *wr = pl->w + 2 * ((long)th * B + b) * S;
              for (long i = 0; i < S; ++i) {
                const double a = xr[2 * i], bb = xr[2 * i + 1], c = wr[2 * i], d = wr[2 * i + 1];
                acc[2 * i] += c * a - d * bb;
                acc[2 * i + 1] += c * bb + d * a;
              }
            }
            for (long i = 0; i < S; ++i) {
              g[2 * (i * RB + rr)] = acc[2 * i];
              g[2 * (i * RB + rr) + 1] = acc[2 * i + 1];
            }
          }
        }
      }
      /* zero the unused rows of a partial block, then the S-point DFT over the batch */
      for (long i = 0; i < S; ++i)
        for (long rr = nj * n_mu; rr < RB; ++rr) g[2 * (i * RB + rr)] = g[2 * (i * RB + rr) + 1] = 0;
      const double *y = port_fft_run(&pl->fS, g, g2, RB);
      const long R0 = J0 * n_mu, nr = nj * n_mu;
      for (long s = 0; s < S; ++s)
        memcpy(pl->at + 2 * (s * Mh + R0), y + 2 * s * RB, sizeof(double) * 2 * nr);
    }
    free(acc);
    free(g);
    free(g2);
  }
}

/* Step 2 + demodulation: four-step FFT of each segment of alpha_tilde, X[s*M + m] =
 * W_inv[m] * y[s][m] for m < M. */
static void port_fft_demod(port_plan *pl, double *X)
{
  const long S = pl->S, M = pl->M, Mh = pl->Mh;
  const int n1 = pl->n1, n2 = pl->n2;
  /* pass 1: columns r2 of each segment (n1 points, stride n2), blocks of PORT_CB columns,
   * then the twiddle w_Mh^{r2 m1}; in place */
  const long ncb = (n2 + PORT_CB - 1) / PORT_CB;
#pragma omp parallel
  {
    double *b0 = (double *)malloc(sizeof(double) * 2 * (size_t)n1 * PORT_CB);
    double *b1 = (double *)malloc(sizeof(double) * 2 * (size_t)n1 * PORT_CB);
#pragma omp for schedule(static) collapse(2)
    for (long s = 0; s < S; ++s)
      for (long cb = 0; cb < ncb; ++cb) {
        double *seg = pl->at + 2 * s * Mh;
        const long c0 = cb * PORT_CB;
        const int nc = c0 + PORT_CB <= n2 ? PORT_CB : (int)(n2 - c0);
        for (long r1 = 0; r1 < n1; ++r1)
          for (int c = 0; c < PORT_CB; ++c) {
            const int cc = c < nc ? c : 0;
            b0[2 * (r1 * PORT_CB + c)] = seg[2 * (r1 * n2 + c0 + cc)];
            b0[2 * (r1 * PORT_CB + c) + 1] = seg[2 * (r1 * n2 + c0 + cc) + 1];
          }
        const double *y = port_fft_run(&pl->f1, b0, b1, PORT_CB);
        for (long m1 = 0; m1 < n1; ++m1)
          for (int c = 0; c < nc; ++c) {
            const long e = (c0 + c) * m1; /* < M_hat */
            const double *th = pl->thi + 2 * (e >> 10), *tl = pl->tlo + 2 * (e & 1023);
            const double tr = th[0] * tl[0] - th[1] * tl[1], ti = th[0] * tl[1] + th[1] * tl[0];
            const double vr = y[2 * (m1 * PORT_CB + c)], vi = y[2 * (m1 * PORT_CB + c) + 1];
            seg[2 * (m1 * n2 + c0 + c)] = vr * tr - vi * ti;
            seg[2 * (m1 * n2 + c0 + c) + 1] = vr * ti + vi * tr;
          }
      }
    free(b0);
    free(b1);
  }
  /* pass 2: rows m1 (n2 contiguous points), blocks of PORT_CB rows; y[m1 + n1 m2] */
  const long nrb = (n1 + PORT_CB - 1) / PORT_CB;
#pragma omp parallel
  {
    double *b0 = (double *)malloc(sizeof(double) * 2 * (size_t)n2 * PORT_CB);
    double *b1 = (double *)malloc(sizeof(double) * 2 * (size_t)n2 * PORT_CB);
#pragma omp for schedule(static) collapse(2)
    for (long s = 0; s < S; ++s)
      for (long rb = 0; rb < nrb; ++rb) {
        const double *seg = pl->at + 2 * s * Mh;
        const long m10 = rb * PORT_CB;
        const int nr = m10 + PORT_CB <= n1 ? PORT_CB : (int)(n1 - m10);
        for (long r2 = 0; r2 < n2; ++r2)
          for (int c = 0; c < PORT_CB; ++c) {
            const int cc = c < nr ? c : 0;
            b0[2 * (r2 * PORT_CB + c)] = seg[2 * ((m10 + cc) * n2 + r2)];
            b0[2 * (r2 * PORT_CB + c) + 1] = seg[2 * ((m10 + cc) * n2 + r2) + 1];
          }
        const double *y = port_fft_run(&pl->f2, b0, b1, PORT_CB);
        for (long m2 = 0; m2 < n2; ++m2)
          for (int c = 0; c < nr; ++c) {
            const long m = m10 + c + (long)n1 * m2;
            if (m >= M) continue;
            const double vr = y[2 * (m2 * PORT_CB + c)], vi = y[2 * (m2 * PORT_CB + c) + 1];
            const double wr = pl->winv[2 * m], wi = pl->winv[2 * m + 1];
            X[2 * (s * M + m)] = wr * vr - wi * vi;
            X[2 * (s * M + m) + 1] = wr * vi + wi * vr;
          }
      }
    free(b0);
    free(b1);
  }
}

/* Step 1 alone, for profiling */
void port_filter_only(void *h, const double *x) { port_filter((port_plan *)h, x); }

/* compute_soi (non-VLC) in place on x (N complex128 values) */
void port_compute(void *h, double *x)
{
  port_plan *pl = (port_plan *)h;
  port_filter(pl, x);
  port_fft_demod(pl, x);
}

int port_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
