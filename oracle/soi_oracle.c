/*
 * soi_oracle.c -- CPU restatement of the SOI FFT hot path (TEST INFRASTRUCTURE ONLY).
 *
 * This file is the checker, never the product: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product path (soi_fft_amd) never links it.
 *
 * It restates, in plain C99 + OpenMP, what the reference computes for compute_soi with
 * P virtual ranks on one address space (global indices, rank = 0 owns everything):
 *
 *   orc_w_table        pfft.c:42-66 (w_f) + pfft.c:190-205 (fill), long-double windows
 *   orc_w_inv_table    pfft.c:68-82 (W_inv_f) + pfft.c:207-210
 *   orc_filter_subsample  parallel_filter_subsampling.cpp:52-527 (conv 167-287, S-DFT 297-311,
 *                      transpose 387-395, tail rows 403-446, ghost rows 496-524)
 *   orc_fft            the MKL DFTI forward, in-place, unscaled, sign -1 transform that
 *                      pfft.c:224-229 plans and pfft.c:642 runs (mixed radix 2/3/4/5 + generic)
 *   orc_demodulate     pfft.c:649-693
 *   orc_compute_soi    pfft.c:276-721 (non-VLC and VLC branches on one address space)
 *   orc_max_exponent / orc_compress / orc_decompress   compress.c:11-544
 *   orc_vlc_plan       pfft.c:390-440 (per-segment send length and load balance)
 *
 * Pinning (what fixes each function independently of this file): see tests/test_oracle.py.
 *   windows     closed forms evaluated in numpy/mpmath-free float64 + symmetry checks
 *   filter      brute-force long-double matrix-vector product (orc_filter_brute)
 *   fft         brute-force long-double DFT for small L, numpy.fft for large L
 *   compute_soi numpy.fft.fft (exact DFT) within the SOI truncation error, closed forms of
 *               input kinds 0 and 1 (input.c:218-231, 256-281)
 *   VLC         hand-derived bit layouts, round-trip error bounds, the compress.c:546-597 KAT
 *               (with the floor rounding that KAT was written against)
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  long N;       /* global length */
  int P;        /* ranks (virtual segments per k) */
  int k;        /* segments per rank */
  int n_mu, d_mu;
  int B;
  double tau, sigma;
} orc_params;

static const long double PI_L = 3.141592653589793238462643383279502884197169399375L;
static const double PI_D = 3.14159265358979323846;

#define ORC_S(p) ((long)(p)->k * (p)->P)
#define ORC_M(p) ((p)->N / ORC_S(p))
#define ORC_MHAT(p) ((long)(p)->n_mu * ORC_M(p) / (p)->d_mu)

/* ------------------------------------------------------------------------------------------
 * Windows.  w_f follows pfft.c:42-66 literally: t is formed in double exactly as the
 * reference forms it, the trigonometric parts are evaluated in long double.
 * Layout: w[(theta*B + b)*S + i] (the logical index theta*B*S + j*S + ii of pfft.c:196).
 * ---------------------------------------------------------------------------------------- */
/* 0: t formed in double as the reference does (pfft.c:52); 1: t = exact rational rounded once;
 * 2: reference t in the real window y(t), exact t in the phase e^{i pi t}.  Modes 1 and 2 exist
 * to measure how much of the error budget the reference's rounding of t takes, and where. */
static int g_window_mode = 0;
void orc_set_window_mode(int mode) { g_window_mode = mode; }

static void w_f(const orc_params *p, long idx, double *re, double *im)
{
  long S = ORC_S(p), M = ORC_M(p), Mh = ORC_MHAT(p);
  long theta = idx / ((long)p->B * S);
  long kappa = p->B - p->d_mu;
  long j = idx % ((long)p->B * S);
  double t = ((double)theta / Mh - (double)j / p->N + (double)kappa / (2 * M)) * M;
  /* t = theta*d/n - j/S + kappa/2 = (2 theta d S - 2 j n + kappa n S) / (2 n S) */
  long tnum = 2 * theta * p->d_mu * S - 2 * j * p->n_mu + kappa * p->n_mu * S;
  double t_exact = (double)((long double)tnum / (long double)(2 * p->n_mu * S));
  double t_phase = t;
  if (g_window_mode == 1) t = t_phase = t_exact;
  if (g_window_mode == 2) t_phase = t_exact;
  long double y;
  if (t == 0) {
    y = 1;
  } else {
    y = sinl(PI_L * p->tau * t) / (PI_D * p->tau * t) * exp(-PI_D * PI_D * t * t / p->sigma);
  }
  double yd = (double)y;
  double rr = (double)cosl(PI_L * t_phase), ri = (double)sinl(PI_L * t_phase);
  double mu = (double)p->n_mu / p->d_mu;
  /* r*y/mu in double complex arithmetic (pfft.c:62-65) */
  *re = rr * yd / mu;
  *im = ri * yd / mu;
}

void orc_w_table(const orc_params *p, double *w)
{
  long S = ORC_S(p);
  long n = (long)p->n_mu * p->B * S;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < n; ++idx) w_f(p, idx, &w[2 * idx], &w[2 * idx + 1]);
}

void orc_w_inv_table(const orc_params *p, double *W_inv)
{
  long M = ORC_M(p);
  long kappa = p->B - p->d_mu;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < M; ++i) {
    double t = (double)i / M - 0.5;
    double y = 1 / (2 * p->tau) *
               (erfc(sqrt(p->sigma) * (t - p->tau / 2)) - erfc(sqrt(p->sigma) * (t + p->tau / 2)));
    double delta = (double)kappa / (2 * M);
    double rr = (double)cosl(2 * PI_L * delta * i);
    double ri = -(double)sinl(2 * PI_L * delta * i);
    /* r/y (pfft.c:81): complex divided by a real */
    W_inv[2 * i] = rr / y;
    W_inv[2 * i + 1] = ri / y;
  }
}

/* ------------------------------------------------------------------------------------------
 * Twiddles and FFT.  orc_fft is the transform MKL's DFTI descriptor computes at
 * pfft.c:224-229 / 642 and parallel_filter_subsampling.cpp:308: forward (e^{-2 pi i nk/L}),
 * in place, no scaling.  Recursive mixed radix; twiddles from long double sincos.
 * ---------------------------------------------------------------------------------------- */
static void twiddle(long num, long den, double *c, double *s)
{
  /* e^{-2 pi i num/den} with exact reduction of num/den */
  long r = num % den;
  if (r < 0) r += den;
  long double a = -2 * PI_L * (long double)r / (long double)den;
  *c = (double)cosl(a);
  *s = (double)sinl(a);
}

static int smallest_factor(long L)
{
  static const int pref[] = {4, 2, 3, 5};
  for (int t = 0; t < 4; ++t)
    if (L % pref[t] == 0) return pref[t];
  for (long f = 7; f * f <= L; f += 2)
    if (L % f == 0) return (int)f;
  return (int)L;
}

/* out[k*os] = sum_n in[n*is] e^{-2 pi i n k / L};  scratch needs 2L doubles */
static void fft_rec(long L, const double *in, long is, double *out, long os, double *scratch)
{
  if (L == 1) {
    out[0] = in[0];
    out[1] = in[1];
    return;
  }
  int r = smallest_factor(L);
  long m = L / r;
  if (r == L && L > 5) {
    /* prime length: direct DFT */
    for (long kk = 0; kk < L; ++kk) {
      long double sr = 0, si = 0;
      for (long n = 0; n < L; ++n) {
        double c, s;
        twiddle(n * kk, L, &c, &s);
        double xr = in[2 * n * is], xi = in[2 * n * is + 1];
        sr += (long double)xr * c - (long double)xi * s;
        si += (long double)xr * s + (long double)xi * c;
      }
      out[2 * kk * os] = (double)sr;
      out[2 * kk * os + 1] = (double)si;
    }
    return;
  }
  /* decimation in time: sub-transform q over inputs n = n1*r + q */
  for (int q = 0; q < r; ++q)
    fft_rec(m, in + 2 * q * is, is * r, scratch + 2 * q * m, 1, out);
  for (long k1 = 0; k1 < m; ++k1) {
    double v[2 * 8];
    double *vv = v;
    double *heap = NULL;
    if (r > 8) heap = vv = (double *)malloc(sizeof(double) * 2 * r);
    for (int q = 0; q < r; ++q) {
      double c, s;
      twiddle((long)q * k1, L, &c, &s);
      double xr = scratch[2 * (q * m + k1)], xi = scratch[2 * (q * m + k1) + 1];
      vv[2 * q] = xr * c - xi * s;
      vv[2 * q + 1] = xr * s + xi * c;
    }
    for (int k2 = 0; k2 < r; ++k2) {
      double sr = 0, si = 0;
      for (int q = 0; q < r; ++q) {
        double c, s;
        twiddle((long)q * k2, r, &c, &s);
        sr += vv[2 * q] * c - vv[2 * q + 1] * s;
        si += vv[2 * q] * s + vv[2 * q + 1] * c;
      }
      long kk = k1 + (long)k2 * m;
      out[2 * kk * os] = sr;
      out[2 * kk * os + 1] = si;
    }
    free(heap);
  }
}

void orc_fft(long L, double *data)
{
  double *tmp = (double *)malloc(sizeof(double) * 2 * L);
  double *scr = (double *)malloc(sizeof(double) * 2 * L);
  fft_rec(L, data, 1, tmp, 1, scr);
  memcpy(data, tmp, sizeof(double) * 2 * L);
  free(tmp);
  free(scr);
}

/* batched: n transforms of length L, contiguous */
void orc_fft_batch(long L, long n, double *data)
{
#pragma omp parallel for schedule(dynamic)
  for (long b = 0; b < n; ++b) orc_fft(L, data + 2 * b * L);
}

/* Brute-force DFT with long double accumulation (pins orc_fft). */
void orc_dft_brute(long L, const double *in, double *out)
{
#pragma omp parallel for schedule(static)
  for (long kk = 0; kk < L; ++kk) {
    long double sr = 0, si = 0;
    for (long n = 0; n < L; ++n) {
      long r = (long)(((__int128)n * kk) % L);
      long double a = -2 * PI_L * (long double)r / (long double)L;
      long double c = cosl(a), s = sinl(a);
      sr += in[2 * n] * c - in[2 * n + 1] * s;
      si += in[2 * n] * s + in[2 * n + 1] * c;
    }
    out[2 * kk] = (double)sr;
    out[2 * kk + 1] = (double)si;
  }
}

/* ------------------------------------------------------------------------------------------
 * Step 1: filter and subsample, global form (parallel_filter_subsampling.cpp:52-527 with
 * all P ranks' row ranges side by side; the halo of the last rank wraps to row 0,
 * parallel_filter_subsampling.cpp:63-64, 101-108).
 *   gamma[R][i] = sum_{b<B} w[theta][b][i] * alpha[((J*d_mu+b) mod M)*S + i],  R = J*n_mu+theta
 * accumulated in b order with the product formed first (intrinsic.h:24-32: fmaddsub of the
 * duplicated real/imag window parts, then _MM_ADD).
 *   alpha_tilde[s][R] = sum_i gamma[R][i] e^{-2 pi i i s/S}     (MKL S-point DFT, :308)
 * Output layout [S][M_hat]: segment s, row R (rank p owns R in [p*l,(p+1)*l), :391/437/522).
 * ---------------------------------------------------------------------------------------- */
void orc_filter_subsample(const orc_params *p, const double *w, const double *alpha,
                          double *alpha_tilde)
{
  long S = ORC_S(p), M = ORC_M(p), Mh = ORC_MHAT(p);
  int n_mu = p->n_mu, d_mu = p->d_mu, B = p->B;
  double *tw = (double *)malloc(sizeof(double) * 2 * S);
  for (long m = 0; m < S; ++m) twiddle(m, S, &tw[2 * m], &tw[2 * m + 1]);
#pragma omp parallel
  {
    double *g = (double *)malloc(sizeof(double) * 2 * S);
#pragma omp for schedule(static)
    for (long R = 0; R < Mh; ++R) {
      long J = R / n_mu;
      int theta = (int)(R % n_mu);
      for (long i = 0; i < S; ++i) {
        double ar = 0, ai = 0;
        for (int b = 0; b < B; ++b) {
          long q = (J * d_mu + b) % M;
          const double *x = &alpha[2 * (q * S + i)];
          const double *ww = &w[2 * (((long)theta * B + b) * S + i)];
          double pr = ww[0] * x[0] - ww[1] * x[1];
          double pi = ww[0] * x[1] + ww[1] * x[0];
          ar += pr;
          ai += pi;
        }
        g[2 * i] = ar;
        g[2 * i + 1] = ai;
      }
      for (long s = 0; s < S; ++s) {
        double sr = 0, si = 0;
        for (long i = 0; i < S; ++i) {
          const double *t = &tw[2 * ((i * s) % S)];
          sr += g[2 * i] * t[0] - g[2 * i + 1] * t[1];
          si += g[2 * i] * t[1] + g[2 * i + 1] * t[0];
        }
        alpha_tilde[2 * (s * Mh + R)] = sr;
        alpha_tilde[2 * (s * Mh + R) + 1] = si;
      }
    }
    free(g);
  }
  free(tw);
}

/* Brute force of the same matrix-vector product in long double, straight from the block
 * structure described at parallel_filter_subsampling.cpp:17-47 (each S-by-S block is
 * DFT_S * diag(window)).  Pins orc_filter_subsample on tiny inputs. */
void orc_filter_brute(const orc_params *p, const double *alpha, double *alpha_tilde)
{
  long S = ORC_S(p), M = ORC_M(p), Mh = ORC_MHAT(p);
  long kappa = p->B - p->d_mu;
  long double mu = (long double)p->n_mu / p->d_mu;
  for (long R = 0; R < Mh; ++R) {
    long J = R / p->n_mu, theta = R % p->n_mu;
    for (long s = 0; s < S; ++s) {
      long double sr = 0, si = 0;
      for (int b = 0; b < p->B; ++b) {
        long q = (J * p->d_mu + b) % M;
        for (long i = 0; i < S; ++i) {
          long double t = (long double)theta * p->d_mu / p->n_mu - b - (long double)i / S +
                          (long double)kappa / 2;
          long double y = (t == 0) ? 1.0L
                                   : sinl(PI_L * p->tau * t) / (PI_L * p->tau * t) *
                                         expl(-PI_L * PI_L * t * t / p->sigma);
          long double ph = PI_L * t - 2 * PI_L * (long double)((i * s) % S) / S;
          long double cr = cosl(ph) * y / mu, ci = sinl(ph) * y / mu;
          long double xr = alpha[2 * (q * S + i)], xi = alpha[2 * (q * S + i) + 1];
          sr += cr * xr - ci * xi;
          si += cr * xi + ci * xr;
        }
      }
      alpha_tilde[2 * (s * Mh + R)] = (double)sr;
      alpha_tilde[2 * (s * Mh + R) + 1] = (double)si;
    }
  }
}

/* pfft.c:649-693: X[s*M + m] = W_inv[m] * y[s][m], m < M (complex product) */
void orc_demodulate(const orc_params *p, const double *W_inv, const double *y, double *X)
{
  long S = ORC_S(p), M = ORC_M(p), Mh = ORC_MHAT(p);
#pragma omp parallel for schedule(static)
  for (long s = 0; s < S; ++s)
    for (long m = 0; m < M; ++m) {
      double wr = W_inv[2 * m], wi = W_inv[2 * m + 1];
      double yr = y[2 * (s * Mh + m)], yi = y[2 * (s * Mh + m) + 1];
      X[2 * (s * M + m)] = wr * yr - wi * yi;
      X[2 * (s * M + m) + 1] = wr * yi + wi * yr;
    }
}

/* ------------------------------------------------------------------------------------------
 * VLC codec, compress.c:11-544.  The reference packs with AVX2 in 16 interleaved 32-bit word
 * lanes (lane j carries elements i+j of every 16-element chunk) and a scalar tail stream.
 * round_mode: 0 = the live code (vector body: round-half-even via _mm256_round_pd and
 * _mm256_cvtpd_epi32, compress.c:235-280; scalar tail: C round(), half away from zero,
 * compress.c:335-337), 1 = floor everywhere (the rounding the commented-out known-answer test
 * at compress.c:546-597 was written against).  Only tests use round_mode = 1.
 * ---------------------------------------------------------------------------------------- */
int orc_max_exponent(const double *x, int len)
{
  /* compress.c:46-93: biased exponent field of every element, max, minus 1023.
   * With no element at all the reference returns INT_MIN - 1023 (two's complement wrap). */
  int gmax = INT_MIN;
  for (int i = 0; i < len; ++i) {
    uint64_t u;
    memcpy(&u, &x[i], 8);
    int e = (int)((u >> 52) & 0x7FF);
    if (e > gmax) gmax = e;
  }
  return (int)((unsigned)gmax - 1023u);
}

static inline uint32_t lsr32(uint32_t v, int sh) { return sh >= 32 ? 0u : (v >> sh); }
static inline uint32_t lsl32(uint32_t v, int sh) { return sh >= 32 ? 0u : (v << sh); }

static double rnd_half_even(double v) { return nearbyint(v); } /* default FE_TONEAREST */

int orc_compress_mode(int *out, const double *in, int len, int e_max, int e_max_i, int round_mode)
{
  const int i1Len = 24 - (e_max - e_max_i);            /* compress.c:207 */
  const int i2Len = (55 - (e_max - e_max_i)) < 32 ? (55 - (e_max - e_max_i)) : 32; /* :208 */
  if (i2Len <= 0) return 0;                              /* :210-212 */
  double e1 = ldexp(1.0, 21 - e_max);                    /* pow(2, (52-31)-e_max) */
  double e2 = ldexp(1.0, 52 - e_max);
  double e3 = ldexp(1.0, 31);

  uint32_t cw[16] = {0};
  int occ = 0, outIdx = 0;
  int i = 0;
  for (; i < len / 16 * 16; i += 16) {
    int32_t v1[16], v2[16];
    for (int j = 0; j < 16; ++j) {
      double x = in[i + j];
      double d1 = round_mode ? floor(x * e1) : rnd_half_even(x * e1);
      double d2 = fma(x, e2, -(d1 * e3));                /* _mm256_fmsub_pd, :238 */
      if (round_mode) d2 = floor(d2);
      v1[j] = (int32_t)rnd_half_even(d1);                /* _mm256_cvtpd_epi32 (MXCSR RNE) */
      v2[j] = (int32_t)rnd_half_even(d2);
    }
    /* append8 (compress.c:170-203) for i1 (when i1Len > 0) then i2 */
    for (int pass = 0; pass < 2; ++pass) {
      int nb = pass == 0 ? i1Len : i2Len;
      if (pass == 0 && nb <= 0) continue;
      uint32_t mask = nb >= 32 ? 0xFFFFFFFFu : ((1u << nb) - 1u);
      uint32_t vv[16];
      for (int j = 0; j < 16; ++j) vv[j] = (uint32_t)(pass == 0 ? v1[j] : v2[j]) & mask;
      occ += nb;
      if (occ >= 32) {
        int nb2 = occ - 32;
        for (int j = 0; j < 16; ++j) out[outIdx + j] = (int)(cw[j] | lsr32(vv[j], nb2));
        outIdx += 16;
        uint32_t m2 = nb2 >= 32 ? 0xFFFFFFFFu : ((1u << nb2) - 1u);
        occ = nb2;
        for (int j = 0; j < 16; ++j) cw[j] = lsl32(vv[j] & m2, 32 - occ);
      } else {
        for (int j = 0; j < 16; ++j) cw[j] |= lsl32(vv[j], 32 - occ);
      }
    }
  }
  if (occ) {                                             /* :324-328 flush the partial row */
    for (int j = 0; j < 16; ++j) out[outIdx + j] = (int)cw[j];
    outIdx += 16;
  }

  /* scalar tail, compress.c:330-369, C round() (half away from zero), unfused arithmetic */
  uint32_t cword = 0;
  occ = 0;
  for (; i < len; ++i) {
    double x = in[i];
    double d1 = round_mode ? floor(x * e1) : round(x * e1);
    double prod = x * e2;
    double d2 = round_mode ? floor(prod - d1 * e3) : round(prod - d1 * e3);
    int32_t iv[2] = {(int32_t)d1, (int32_t)d2};
    for (int pass = 0; pass < 2; ++pass) {
      int nb = pass == 0 ? i1Len : i2Len;
      if (pass == 0 && nb <= 0) continue;
      uint32_t v = (uint32_t)iv[pass] & (nb >= 32 ? 0xFFFFFFFFu : ((1u << nb) - 1u));
      occ += nb;
      if (occ >= 32) {
        int nb2 = occ - 32;
        cword |= lsr32(v, nb2);
        out[outIdx++] = (int)cword;
        cword = 0;
        v &= nb2 >= 32 ? 0xFFFFFFFFu : ((1u << nb2) - 1u);
        occ = nb2;
      }
      cword |= lsl32(v, 32 - occ);
    }
  }
  if (occ) out[outIdx++] = (int)cword;
  return outIdx;
}

int orc_compress(int *out, const double *in, int len, int e_max, int e_max_i)
{
  return orc_compress_mode(out, in, len, e_max, e_max_i, 0);
}

/* decompress, compress.c:428-544.  Vector body mirrors extract8 (:395-426).  The tail is the
 * exact inverse of compress's scalar stream (which starts right after the vector rows).  The
 * reference's tail setup (:491-498) starts reading 16 words past the partial row, or one word
 * late after a full row, which does not invert its own compress; that quirk is not restated.
 * When i2Len <= 0 compress sent nothing and the elements decode to 0. */
static inline int32_t sx_field(uint32_t cw, int occ, int nbits)
{
  /* (curWord << (32-occ)) >> (32-nbits), arithmetic */
  uint32_t sh = lsl32(cw, 32 - occ);
  int32_t s = (int32_t)sh;
  int r = 32 - nbits;
  return r >= 32 ? (s < 0 ? -1 : 0) : (s >> r);
}

void orc_decompress(double *out, const int *in, int len, int e_max, int e_max_i)
{
  double e1 = ldexp(1.0, e_max - 21);
  double e2 = ldexp(1.0, e_max - 52);
  const int i1Len = 24 - (e_max - e_max_i);
  const int i2Len = (55 - (e_max - e_max_i)) < 32 ? (55 - (e_max - e_max_i)) : 32;
  if (i2Len <= 0) {
    for (int i = 0; i < len; ++i) out[i] = 0;
    return;
  }
  int i = 0;
  int nrows_vec = 0;
  if (len / 16 > 0) {
    uint32_t cw[16];
    for (int j = 0; j < 16; ++j) cw[j] = (uint32_t)in[j];
    int occ = 32, inIdx = 16;
    for (; i < len / 16 * 16; i += 16) {
      int32_t f[2][16];
      for (int pass = 0; pass < 2; ++pass) {
        int nb = pass == 0 ? i1Len : i2Len;
        if (pass == 0 && nb <= 0) {
          for (int j = 0; j < 16; ++j) f[0][j] = 0;
          continue;
        }
        for (int j = 0; j < 16; ++j) f[pass][j] = sx_field(cw[j], occ, nb);
        if (nb <= occ) {
          occ -= nb;
          if (occ == 0) {
            for (int j = 0; j < 16; ++j) cw[j] = (uint32_t)in[inIdx + j];
            inIdx += 16;
            occ = 32;
          }
        } else {
          int need = nb - occ;
          for (int j = 0; j < 16; ++j) cw[j] = (uint32_t)in[inIdx + j];
          for (int j = 0; j < 16; ++j) f[pass][j] |= (int32_t)lsr32(cw[j], 32 - need);
          inIdx += 16;
          occ = 32 - need;
        }
      }
      for (int j = 0; j < 16; ++j)
        out[i + j] = fma((double)f[0][j], e1, (double)f[1][j] * e2); /* :472-483 */
    }
    /* rows the vector body occupies: ceil(bits per lane / 32) */
    long bits = (long)(len / 16) * ((i1Len > 0 ? i1Len : 0) + i2Len);
    nrows_vec = (int)((bits + 31) / 32);
  }
  int inIdx = nrows_vec * 16;
  uint32_t cword = 0;
  int occ = 0;
  for (; i < len; ++i) {
    int32_t f[2] = {0, 0};
    for (int pass = 0; pass < 2; ++pass) {
      int nb = pass == 0 ? i1Len : i2Len;
      if (pass == 0 && nb <= 0) continue;
      if (occ == 0) {
        cword = (uint32_t)in[inIdx++];
        occ = 32;
      }
      int32_t v = sx_field(cword, occ, nb);
      if (nb <= occ) {
        occ -= nb;
      } else {
        int need = nb - occ;
        cword = (uint32_t)in[inIdx++];
        v |= (int32_t)lsr32(cword, 32 - need);
        occ = 32 - need;
      }
      f[pass] = v;
    }
    out[i] = (double)f[0] * e1 + (double)f[1] * e2;     /* :529-531 */
  }
}

/* pfft.c:390-440: per-segment send count and cost-balanced segment boundaries.
 * gmax[s] = per-segment max exponent after the Allreduce, totalMax their max.
 * Outputs sCnts[S] (in doubles, (sendLen+1)/2) and segBound[P+1]. */
void orc_vlc_plan(int P, int k, long l, long M_hat, double comm_to_comp, const int *gmax,
                  int *sCnts, int *segBound)
{
  int S = P * k;
  int totalMax = INT_MIN;
  for (int s = 0; s < S; ++s)
    if (gmax[s] > totalMax) totalMax = gmax[s];
  double *pcs = (double *)malloc(sizeof(double) * (S + 1));
  pcs[0] = 0;
  for (int p = 0; p < P; ++p)
    for (int ik = 0; ik < k; ++ik) {
      int s = p * k + ik;
      int bitsPerInput = 52 + 3 - (totalMax - gmax[s]);
      if (bitsPerInput > 31) ++bitsPerInput;
      int lenPerSimdLane = (int)((l * 2 / 16 * bitsPerInput + 31) / 32);
      int remainder = (int)((l * 2 % 16 * bitsPerInput + 31) / 32);
      int sendLen = lenPerSimdLane * 16 + remainder;
      if (bitsPerInput <= 0) sendLen = 0;
      sCnts[s] = (sendLen + 1) / 2;
      pcs[s + 1] = pcs[s] + comm_to_comp * sCnts[s] / 2 / M_hat + 1;
    }
  for (int p = 0; p <= P; ++p) segBound[p] = p * k;
  double costPerRank = pcs[S] / P;
  int s = 1;
  for (int p = 1; p < P;) {
    if (pcs[s] >= p * costPerRank) {
      if (pcs[s] - p * costPerRank > p * costPerRank - pcs[s - 1])
        segBound[p] = s - 1;
      else
        segBound[p] = s;
      ++p;
    } else {
      ++s;
    }
  }
  free(pcs);
}

/* ------------------------------------------------------------------------------------------
 * compute_soi on one address space (pfft.c:276-721).  alpha (2N doubles) is overwritten with
 * the N-point DFT in natural order: X[s*M + m].  With use_vlc, every (segment, rank) chunk of
 * l elements is compressed with (totalMax, gmax[s]) and decompressed before the M_hat-point FFT
 * (pfft.c:316-353, 476-543, 618-632).  seg_bound (P+1 ints) receives segmentBoundaries.
 * ---------------------------------------------------------------------------------------- */
int orc_compute_soi(const orc_params *p, double *alpha, int use_vlc, double comm_to_comp,
                    int *seg_bound)
{
  long S = ORC_S(p), M = ORC_M(p), Mh = ORC_MHAT(p);
  long l = Mh / p->P;
  double *w = (double *)malloc(sizeof(double) * 2 * p->n_mu * p->B * S);
  double *W_inv = (double *)malloc(sizeof(double) * 2 * M);
  double *at = (double *)malloc(sizeof(double) * 2 * S * Mh);
  if (!w || !W_inv || !at) return -1;
  orc_w_table(p, w);
  orc_w_inv_table(p, W_inv);
  orc_filter_subsample(p, w, alpha, at);
  for (int q = 0; q <= p->P; ++q)
    if (seg_bound) seg_bound[q] = q * p->k;
  if (use_vlc) {
    int *gmax = (int *)malloc(sizeof(int) * S);
    int *sc = (int *)malloc(sizeof(int) * S);
    int *buf = (int *)malloc(sizeof(int) * 4 * l + 64);
    int tmax = INT_MIN;
    for (long s = 0; s < S; ++s) {
      int m = INT_MIN;
      for (int q = 0; q < p->P; ++q) {
        int e = orc_max_exponent(at + 2 * (s * Mh + q * l), (int)(2 * l));
        if (e > m) m = e;
      }
      gmax[s] = m;
      if (m > tmax) tmax = m;
    }
    int *sb = (int *)malloc(sizeof(int) * (p->P + 1));
    orc_vlc_plan(p->P, p->k, l, Mh, comm_to_comp, gmax, sc, sb);
    if (seg_bound) memcpy(seg_bound, sb, sizeof(int) * (p->P + 1));
    for (long s = 0; s < S; ++s)
      for (int q = 0; q < p->P; ++q) {
        double *chunk = at + 2 * (s * Mh + q * l);
        orc_compress(buf, chunk, (int)(2 * l), tmax, gmax[s]);
        orc_decompress(chunk, buf, (int)(2 * l), tmax, gmax[s]);
      }
    free(gmax);
    free(sc);
    free(buf);
    free(sb);
  }
  orc_fft_batch(Mh, S, at);
  orc_demodulate(p, W_inv, at, alpha);
  free(w);
  free(W_inv);
  free(at);
  return 0;
}

/* ------------------------------------------------------------------------------------------
 * Input generators shared by tests and the CPU baseline (the product has its own GPU
 * generator; these are for checking).  splitmix64 counter generator.
 * ---------------------------------------------------------------------------------------- */
static inline uint64_t splitmix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void orc_uniform_input(long N, uint64_t seed, double *x)
{
#pragma omp parallel for schedule(static)
  for (long n = 0; n < 2 * N; ++n) {
    uint64_t u = splitmix64(seed ^ (uint64_t)n * 0xD1B54A32D192ED03ull);
    x[n] = 2.0 * ((double)(u >> 11) * 0x1.0p-53) - 1.0;
  }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
