/*
 * soi_test.c -- command-line driver for the SOI FFT on one GPU, with the options and the
 * "key<TAB>value" output of the reference's test driver (test.c:32-138 options, 197-472 main
 * loop) and its accuracy metrics (input.c:218-281 input kinds and exact transforms, 374-425
 * compute_snr / compute_normalized_inf_norm).
 *
 *   soi_test [--P P] [--k_min k] [--k_max K] [--input_min i] [--input_max I] [--n_mu n]
 *            [--d_mu d] [--tau t] [--sigma s] [--B B] [--vlc] [--comm_to_comp_ratio r]
 *            [--no_mkl] [--no_soi] [--no_snr] [--fp32] [--in_file f] [--mkl_out_file f]
 *            [--soi_out_file f] N            (N may be written a^b, e.g. 2^27)
 *
 * Differences from the reference driver, which runs one MPI rank per process:
 *  - P is an option (virtual ranks on one GPU), not the MPI world size;
 *  - the non-SOI comparison FFT (the reference's MKL / MKL-CDFT leg, keys *_mkl) is rocFFT's
 *    FP64 (or FP32) transform on the same GPU, reported under the keys *_ref;
 *  - input kinds 0 (rectangle), 1 (e^{i n theta}) and 3 (Dirichlet kernel) are supported.
 *    Kind 2 needs the reference's particle table and VML; kinds 4 and 5 read data files that
 *    do not ship; they are rejected.
 *  - kind 3's expected output is the exact transform N/W on the rectangle (the reference
 *    compares against 1/W there);
 *  - times are wall-clock seconds around compute_soi on a device-resident buffer.
 */
#define _GNU_SOURCE
#include <getopt.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <hip/hip_runtime_api.h>
#include <rocfft/rocfft.h>

#include "../../include/soi.h"

typedef struct {
  double re, im;
} c64;

typedef struct {
  int P, k_min, k_max, input_min, input_max;
  int no_ref, no_soi, no_snr, fp32;
  const char *in_file, *ref_out_file, *soi_out_file;
} options;

static double now(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void die(const char *what)
{
  fprintf(stderr, "soi_test: %s\n", what);
  exit(1);
}

static void hip_ok(hipError_t e, const char *what)
{
  if (e != hipSuccess) {
    fprintf(stderr, "soi_test: %s: %s\n", what, hipGetErrorString(e));
    exit(1);
  }
}

static void soi_ok(const char *what)
{
  if (soi_last_status()) {
    fprintf(stderr, "soi_test: %s: %s\n", what, soi_last_error());
    exit(1);
  }
}

/* ---- inputs and accuracy metrics: the library's harness entry points ---------------------- */

static int kind_supported(int kind) { return kind == 0 || kind == 1 || kind == 3; }

static void populate(c64 *x, size_t N, int kind)
{
  populate_input((cfft_complex_t *)x, N, 0, N, kind);
  soi_ok("populate_input");
}

/* SNR and normalised infinity-norm error of out against the exact transform.  For kind 3 the
 * library's reference_output is the reference's 1/W rectangle, 1/N of the exact transform, so
 * out is compared after scaling by 1/N (both metrics are ratios; the scaling leaves them
 * unchanged up to rounding).  out is overwritten in that case. */
static void metrics(c64 *out, size_t N, int kind, double *snr, double *inf_err)
{
  if (kind == 3) {
    const double s = 1.0 / (double)N;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; ++i) out[i].re *= s, out[i].im *= s;
  }
  *snr = compute_snr((cfft_complex_t *)out, N, 0, N, kind, NULL);
  *inf_err = compute_normalized_inf_norm((cfft_complex_t *)out, N, 0, N, kind);
  soi_ok("compute_snr");
}

static void write_file(const char *name, const c64 *x, size_t N)
{
  if (!name) return;
  FILE *fp = fopen(name, "w");
  if (!fp) die("cannot open output file");
  for (size_t i = 0; i < N; ++i) fprintf(fp, "%g %g\n", x[i].re, x[i].im);
  fclose(fp);
}

/* host c64 <-> device buffer of complex128 or complex64 */
static void to_device(void *dev, const c64 *h, size_t N, int fp32, void *stage)
{
  if (!fp32) {
    hip_ok(hipMemcpy(dev, h, N * sizeof(c64), hipMemcpyHostToDevice), "copy in");
    return;
  }
  float *s = (float *)stage;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) s[2 * i] = (float)h[i].re, s[2 * i + 1] = (float)h[i].im;
  hip_ok(hipMemcpy(dev, s, N * 2 * sizeof(float), hipMemcpyHostToDevice), "copy in");
}

static void from_device(c64 *h, const void *dev, size_t N, int fp32, void *stage)
{
  if (!fp32) {
    hip_ok(hipMemcpy(h, dev, N * sizeof(c64), hipMemcpyDeviceToHost), "copy out");
    return;
  }
  float *s = (float *)stage;
  hip_ok(hipMemcpy(s, dev, N * 2 * sizeof(float), hipMemcpyDeviceToHost), "copy out");
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; ++i) h[i].re = s[2 * i], h[i].im = s[2 * i + 1];
}

/* ---- options (test.c:32-138) --------------------------------------------------------------- */

static size_t parse_n(const char *s)
{
  const char *caret = strchr(s, '^');
  if (!caret) return (size_t)atol(s);
  return (size_t)llround(pow((double)atol(s), (double)atol(caret + 1)));
}

static void usage(const char *argv0)
{
  fprintf(stderr,
          "usage: %s [--P P] [--k_min k] [--k_max K] [--input_min i] [--input_max I] "
          "[--n_mu n] [--d_mu d] [--tau t] [--sigma s] [--B B] [--vlc] "
          "[--comm_to_comp_ratio r] [--no_mkl] [--no_soi] [--no_snr] [--fp32] [--in_file f] "
          "[--mkl_out_file f] [--soi_out_file f] N\n",
          argv0);
}

static options parse(int argc, char **argv, soi_desc_t *d)
{
  options o;
  memset(&o, 0, sizeof(o));
  o.P = 8;
  o.k_min = o.k_max = 8;
  o.input_min = o.input_max = 1;
  static struct option lo[] = {
      {"k_min", required_argument, 0, 'k'},        {"k_max", required_argument, 0, 'K'},
      {"input_min", required_argument, 0, 'i'},    {"input_max", required_argument, 0, 'I'},
      {"no_mkl", no_argument, 0, 'o'},             {"no_soi", no_argument, 0, 'O'},
      {"no_snr", no_argument, 0, 'c'},             {"n_mu", required_argument, 0, 'n'},
      {"d_mu", required_argument, 0, 'd'},         {"tau", required_argument, 0, 'a'},
      {"sigma", required_argument, 0, 'g'},        {"B", required_argument, 0, 'B'},
      {"in_file", required_argument, 0, 'x'},      {"mkl_out_file", required_argument, 0, 'm'},
      {"soi_out_file", required_argument, 0, 's'}, {"vlc", no_argument, 0, 'v'},
      {"comm_to_comp_ratio", required_argument, 0, 'r'},
      {"P", required_argument, 0, 'P'},            {"fp32", no_argument, 0, 'F'},
      {0, 0, 0, 0}};
  for (;;) {
    int idx = 0;
    int c = getopt_long(argc, argv, "k:K:n:d:B:m:s:P:", lo, &idx);
    if (c == -1) break;
    switch (c) {
      case 'k': o.k_min = atoi(optarg); break;
      case 'K': o.k_max = atoi(optarg); break;
      case 'i': o.input_min = atoi(optarg); break;
      case 'I': o.input_max = atoi(optarg); break;
      case 'o': o.no_ref = 1; break;
      case 'O': o.no_soi = 1; break;
      case 'c': o.no_snr = 1; break;
      case 'n': d->n_mu = atoi(optarg); break;
      case 'd': d->d_mu = atoi(optarg); break;
      case 'a': d->tau = atof(optarg) / 1024; break;   /* test.c:95: tau in 1/1024 units */
      case 'g': d->sigma = atof(optarg); break;
      case 'B': d->B = (cfft_size_t)atoi(optarg); break;
      case 'x': o.in_file = optarg; break;
      case 'm': o.ref_out_file = optarg; break;
      case 's': o.soi_out_file = optarg; break;
      case 'v': d->use_vlc = 1; break;
      case 'r': d->comm_to_comp_cost_ratio = atof(optarg); break;
      case 'P': o.P = atoi(optarg); break;
      case 'F': o.fp32 = 1; break;
      default: usage(argv[0]); exit(1);
    }
  }
  if (optind >= argc) {
    usage(argv[0]);
    exit(1);
  }
  d->N = parse_n(argv[optind]);
  d->P = o.P;
  d->rank = 0;
  if (o.P < 1 || o.k_min < 1 || o.k_max < o.k_min) die("need P >= 1 and 1 <= k_min <= k_max");
  /* test.c:130 */
  if (d->N == 0 || d->N % ((size_t)d->d_mu * o.k_max * o.P * 16) != 0) {
    fprintf(stderr, "(d_mu=%d)*(P=%d)*(k=%d)*16 must divide N\n", d->d_mu, o.P, o.k_max);
    exit(1);
  }
  for (int kind = o.input_min; kind <= o.input_max; ++kind)
    if (!kind_supported(kind)) {
      fprintf(stderr, "input kind %d is not supported (kinds 0, 1 and 3 are)\n", kind);
      exit(1);
    }
  d->precision = o.fp32 ? SOI_PREC_FP32 : SOI_PREC_FP64;
  return o;
}

int main(int argc, char **argv)
{
  soi_desc_t d;
  set_default_soi_descriptor(&d);
  options o = parse(argc, argv, &d);
  const size_t N = d.N;
  printf("P = %d, N = %zu, n_mu = %d, d_mu = %d, B = %zu, sigma = %f\n", d.P, N, d.n_mu, d.d_mu,
         d.B, d.sigma);
  const double flop = 5. * N * log2((double)N);
  const size_t esz = o.fp32 ? 2 * sizeof(float) : sizeof(c64);

  c64 *h = (c64 *)malloc(N * sizeof(c64));
  void *stage = o.fp32 ? malloc(N * 2 * sizeof(float)) : NULL;
  if (!h || (o.fp32 && !stage)) die("host allocation failed");
  void *x = NULL;
  hip_ok(hipMalloc(&x, N * esz), "device allocation");

  rocfft_plan plan = NULL;
  rocfft_execution_info info = NULL;
  void *work = NULL;
  if (!o.no_ref) {
    if (rocfft_setup() != rocfft_status_success) die("rocfft_setup failed");
    size_t len = N, wsz = 0;
    if (rocfft_plan_create(&plan, rocfft_placement_inplace, rocfft_transform_type_complex_forward,
                           o.fp32 ? rocfft_precision_single : rocfft_precision_double, 1, &len, 1,
                           NULL) != rocfft_status_success)
      die("rocfft_plan_create failed");
    rocfft_plan_get_work_buffer_size(plan, &wsz);
    rocfft_execution_info_create(&info);
    if (wsz) {
      hip_ok(hipMalloc(&work, wsz), "rocFFT work buffer");
      rocfft_execution_info_set_work_buffer(info, work, wsz);
    }
  }

  const int REPEAT = 4; /* test.c:217 */
  for (int iter = 0; iter < REPEAT; ++iter) {
    for (int kind = o.input_min; kind <= o.input_max; ++kind) {
      if (!o.no_ref) {
        populate(h, N, kind);
        write_file(o.in_file, h, N);
        to_device(x, h, N, o.fp32, stage);
        hip_ok(hipDeviceSynchronize(), "sync");
        double t = now();
        void *bufs[1] = {x};
        if (rocfft_execute(plan, bufs, NULL, info) != rocfft_status_success)
          die("rocfft_execute failed");
        hip_ok(hipDeviceSynchronize(), "sync");
        t = now() - t;
        printf("time_ref\t%f\n", t);
        printf("flops_ref\t%f\n", flop / t / 1e9);
        from_device(h, x, N, o.fp32, stage);
        write_file(o.ref_out_file, h, N);
        if (!o.no_snr) {
          double snr, inf;
          metrics(h, N, kind, &snr, &inf);
          printf("snr_ref%d\t%f\n", kind, snr);
          printf("max_err_ref%d\t%e\n", kind, inf);
        }
      }
      for (int k = o.k_min; k <= o.k_max && !o.no_soi; k *= 2) {
        init_soi_descriptor(&d, MPI_COMM_WORLD, (cfft_size_t)k);
        soi_ok("init_soi_descriptor");
        populate(h, N, kind);
        to_device(x, h, N, o.fp32, stage);
        hip_ok(hipDeviceSynchronize(), "sync");
        double t = now();
        compute_soi(&d, (cfft_complex_t *)x);
        soi_ok("compute_soi");
        t = now() - t;
        printf("time_soi_%d\t%f\n", k, t);
        printf("flops_soi_%d\t%f\n", k, flop / t / 1e9);
        free_soi_descriptor(&d);
        soi_ok("free_soi_descriptor");
        from_device(h, x, N, o.fp32, stage);
        write_file(o.soi_out_file, h, N);
        if (!o.no_snr) {
          double snr, inf;
          metrics(h, N, kind, &snr, &inf);
          printf("snr_soi%d_%d\t%f\n", kind, k, snr);
          printf("max_err_soi%d_%d\t%e\n", kind, k, inf);
        }
        fflush(stdout);
      }
    }
  }
  if (plan) rocfft_plan_destroy(plan);
  if (info) rocfft_execution_info_destroy(info);
  if (work) (void)hipFree(work);
  if (!o.no_ref) rocfft_cleanup();
  (void)hipFree(x);
  free(stage);
  free(h);
  return 0;
}
