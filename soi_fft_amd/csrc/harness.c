/* harness.c -- the reference's test-harness entry points that soi.h exports for harness
 * parity (reference soi.h:282-288, 306, 314-315), as host C on this library's side:
 *
 *   populate_input               input kinds 0, 1 and 3 (input.c:218-281)
 *   reference_output             their closed-form transforms (input.c:256-290)
 *   compute_snr                  10 log10(sum |ref|^2 / sum |out - ref|^2) (input.c:292-393)
 *   compute_normalized_inf_norm  sqrt(max |out - ref|^2 / max |ref|^2) (input.c:395-425)
 *   get_cpu_freq                 TSC ticks per second (cpu_freq.c:8-23)
 *   load_imbalance_times         per-thread timing counters (pfft.c:270)
 *
 * Buffers are host memory, as in the reference.  Kinds 2, 4 and 5 need MKL's VML and CDFT or
 * data files that do not ship; they set SOI_ERR_ARGS (soi_last_error) where the reference
 * asserts, and leave the buffer untouched (compute_* return NaN).
 *
 * On one GPU the whole transform is one "rank": offset and localLen select a window of the
 * global index range exactly as for a reference rank, so a single call with offset 0 and
 * localLen = globalLen covers everything.  The element loops are OpenMP-parallel, as the
 * reference's are (input.c:218-231, 374-424).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <time.h>
#include <x86intrin.h>

#include "../../include/soi.h"

/* error reporting of the library (pfft.c) */
void soi_internal_set_error(int status, const char *msg);

static const long double PI_L = 3.141592653589793238462643383279502884197169399375L;
static const long double THETA = 1;   /* input.c:19 */
static const long INV_REL_W = 16;     /* input.c:21 */

static int kind_ok(int kind)
{
  soi_internal_set_error(0, "");
  if (kind == 0 || kind == 1 || kind == 3) return 1;
  soi_internal_set_error(SOI_ERR_ARGS, kind == 2 || kind == 4 || kind == 5
                                           ? "input kind needs MKL VML/CDFT or a data file"
                                           : "unknown input kind");
  return 0;
}

static void put(cfft_complex_t *z, double re, double im)
{
  double v[2] = {re, im};
  memcpy(z, v, sizeof v);
}

static void get(const cfft_complex_t *z, double *re, double *im)
{
  double v[2];
  memcpy(v, z, sizeof v);
  *re = v[0];
  *im = v[1];
}

/* sin(pi W idx/N) / (W sin(pi idx/N)), 1 at idx = 0, with the range reduction of idx*W/N that
 * input.c:238-245 and 270-276 use */
static double dirichlet(size_t idx, size_t N)
{
  if (idx == 0) return 1;
  const long W = (long)N / INV_REL_W + 1;
  const double A = (double)idx * (double)W / (double)N;
  const long L = (long)(A + .5);
  const double f = ((double)idx * (double)W - (double)L * (double)N) / (double)N;
  const long double num = (L & 1) ? -sinl(PI_L * f) : sinl(PI_L * f);
  const long double den =
      (2 * idx <= N) ? W * sinl(PI_L * idx / N) : W * sinl(PI_L * (N - idx) / N);
  return (double)(num / den);
}

static double rect(size_t idx, size_t N)
{
  const long W = (long)N / INV_REL_W + 1;
  return (2 * idx < (size_t)W || 2 * (N - idx) < (size_t)W) ? 1. / W : 0;
}

void populate_input(cfft_complex_t *input, size_t localLen, size_t offset, size_t globalLen,
                    int kind)
{
  if (!input || !kind_ok(kind)) return;
#pragma omp parallel for schedule(static)
  for (size_t j = 0; j < localLen; ++j) {
    const size_t idx = j + offset;
    if (kind == 1) { /* e^{i idx theta} */
      const long double a = (long double)idx * THETA;
      put(input + j, (double)cosl(a), (double)sinl(a));
    } else if (kind == 0) { /* rectangle of relative width 1/16 */
      put(input + j, rect(idx, globalLen), 0);
    } else { /* kind 3: the Dirichlet kernel */
      put(input + j, dirichlet(idx, globalLen), 0);
    }
  }
}

/* expected output of a supported kind (kind_ok checked by the caller) */
static cfft_complex_t expected(size_t idx, size_t globalLen, int kind)
{
  cfft_complex_t r;
  put(&r, 0, 0);
  if (kind == 1) { /* (1 - e^{i N theta}) / (1 - e^{i (theta - 2 pi idx/N)}) */
    const long double a = THETA * globalLen;
    const long double nr = 1 - cosl(a), ni = -sinl(a);
    const long double ph = THETA - 2 * PI_L * idx / globalLen;
    const long double s = sinl(ph / 2);
    const long double dr = 2 * s * s, di = -sinl(ph);
    const long double dd = dr * dr + di * di;
    put(&r, (double)((nr * dr + ni * di) / dd), (double)((ni * dr - nr * di) / dd));
  } else if (kind == 0) {
    put(&r, dirichlet(idx, globalLen), 0);
  } else {
    put(&r, rect(idx, globalLen), 0);
  }
  return r;
}

/* Kind 3 returns the reference's value (input.c:284-286), the rectangle of height 1/W; the
 * exact forward DFT of the kind-3 input is N times that. */
cfft_complex_t reference_output(size_t idx, size_t globalLen, int kind, size_t offset)
{
  (void)offset;
  if (!kind_ok(kind)) {
    cfft_complex_t r;
    put(&r, NAN, NAN);
    return r;
  }
  return expected(idx, globalLen, kind);
}

static void errors(const cfft_complex_t *output, size_t localLen, size_t offset, size_t N,
                   int kind, double *perr, double *psig, double *merr, double *msig)
{
  double pe = 0, ps = 0, me = 0, ms = 0;
#pragma omp parallel for schedule(static) reduction(+ : pe, ps) reduction(max : me, ms)
  for (size_t j = 0; j < localLen; ++j) {
    const cfft_complex_t r = expected(j + offset, N, kind);
    double rr, ri, orr, oi;
    get(&r, &rr, &ri);
    get(output + j, &orr, &oi);
    const double dr = orr - rr, di = oi - ri;
    const double e = dr * dr + di * di, s = rr * rr + ri * ri;
    pe += e;
    ps += s;
    if (e > me) me = e;
    if (s > ms) ms = s;
  }
  *perr = pe;
  *psig = ps;
  *merr = me;
  *msig = ms;
}

double compute_snr(cfft_complex_t *output, size_t localLen, size_t offset, size_t globalLen,
                   int kind, soi_desc_t *d)
{
  (void)d;
  if (!output || !kind_ok(kind)) return NAN;
  double perr, psig, merr, msig;
  errors(output, localLen, offset, globalLen, kind, &perr, &psig, &merr, &msig);
  return 10 * log10(psig / perr);
}

double compute_normalized_inf_norm(cfft_complex_t *output, size_t localLen, size_t offset,
                                   size_t globalLen, int kind)
{
  if (!output || !kind_ok(kind)) return NAN;
  double perr, psig, merr, msig;
  errors(output, localLen, offset, globalLen, kind, &perr, &psig, &merr, &msig);
  return sqrt(merr / msig);
}

/* TSC ticks per second, measured once against the monotonic clock over about 10 ms */
double get_cpu_freq(void)
{
  static double freq = 0;
  if (freq == 0) {
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    const unsigned long long c0 = __rdtsc();
    double dt = 0;
    do {
      clock_gettime(CLOCK_MONOTONIC, &t1);
      dt = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    } while (dt < 0.01);
    freq = (double)(__rdtsc() - c0) / dt;
  }
  return freq;
}

/* The reference accumulates per-OpenMP-thread decompression and demodulation cycles here.  The
 * GPU path runs no such host threads: the counters exist for link parity and stay zero. */
unsigned long long load_imbalance_times[SOI_MAX_THREADS];
