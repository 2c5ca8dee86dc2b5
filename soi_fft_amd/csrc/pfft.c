/*
 * pfft.c -- C host layer of the MI355X SOI FFT: the soi.h descriptor API over the HIP kernels.
 *
 * Reference counterparts (IntelLabs/SOI_FFT):
 *   set_default_soi_descriptor  pfft.c:84-138
 *   init_soi_descriptor         pfft.c:140-232
 *   free_soi_descriptor         pfft.c:234-267
 *   compute_soi                 pfft.c:276-721
 *   parallel_filter_subsampling parallel_filter_subsampling.cpp:534-550
 * On one GPU the P ranks are P virtual ranks: the halo exchange becomes circular indexing in
 * the filter kernel, the all-to-all becomes the segment-major layout of alpha_tilde, and the
 * VLC branch compresses and decompresses every (segment, rank) chunk on the device.
 */
#define _GNU_SOURCE
#include <limits.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/soi.h"
#include "../../include/compress.h"
#include "soi_kernels.h"

/* ------------------------------------------------------------------------------------------ */
/* error state                                                                                 */
/* ------------------------------------------------------------------------------------------ */
static __thread char g_err[512];
static __thread int g_status;

static void set_err(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  g_status = code;
}
static void clear_err(void)
{
  g_err[0] = 0;
  g_status = 0;
}
const char *soi_last_error(void) { return g_err; }
int soi_last_status(void) { return g_status; }
/* for the library's other host files (harness.c): status 0 clears */
void soi_internal_set_error(int status, const char *msg)
{
  if (!status) clear_err();
  else set_err(status, "%s", msg);
}

#define HIPCHK(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      set_err(SOI_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),      \
              __FILE__, __LINE__);                                                     \
      goto fail;                                                                       \
    }                                                                                  \
  } while (0)

#define KCHK(call)                                                                     \
  do {                                                                                 \
    int e_ = (call);                                                                   \
    if (e_ != 0) {                                                                     \
      set_err(SOI_ERR_HIP, "%s failed: %s (%s:%d)", #call,                             \
              hipGetErrorString((hipError_t)e_), __FILE__, __LINE__);                  \
      goto fail;                                                                       \
    }                                                                                  \
  } while (0)

/* ------------------------------------------------------------------------------------------ */
/* plan                                                                                        */
/* ------------------------------------------------------------------------------------------ */
#define NSTAGE 4
#define NTSLOT 64 /* timed calls kept for soi_stage_times */
#define SOI_MAX_FFT_GROUPS 32
typedef struct {
  int device;
  hipStream_t stream;
  long N, S, M, Mh, l;
  int P, k, n_mu, d_mu, B;
  /* filter */
  SoiFilterArgs fa;
  int ref_windows;
  void *taps;
  soi_c64_t *post, *twS;
  /* FP32 form: copies of the FP64 tables rounded to FP32 (W_inv stays FP64) */
  int f32;
  soi_c32_t *taps32, *post32, *twS32, *tw1_32, *tw2_32, *thi32, *tlo32;
  void *wt; /* complex64 MFMA filter: bf16 tap pieces */
  /* fft */
  int fft_single, single_count, fft_reg;
  int n1, n2, Tc, Tr, B1, lb1, clay;
  /* register-FFT alternatives (column tile, column variant, row tile) and the one in use */
  int nvar, var;
  int vtab[16][3];
  /* pass 1 writes 8-row tiles into tbuf (S*M_hat values) and pass 2 reads them (register
   * path).  A buffer of the plan's own: a caller's gamma_tilde may be sized as the
   * reference's per-rank work buffer. */
  int tiled;
  void *tbuf;
  soi_c64_t *tw1, *tw2, *thi, *tlo;
  /* buffers */
  soi_c64_t *io;        /* staging for host-pointer inputs */
  /* the device buffers the kernels use, kept here as well as in the descriptor fields: a
   * reference-style caller may null or overwrite d->alpha_tilde, d->w and d->W_inv between
   * calls (test.c:406-420), and the plan still reaches and frees what it owns */
  void *at;             /* alpha_tilde: the caller's preallocated buffer or owns_at */
  int owns_at;
  soi_c64_t *wtab, *winv;
  int *delta;
  /* vlc: per-chunk exponents, segment exponents and word counts on the device; the plan of
   * k_vlc_plan ([total, seg_max[S], send_counts[S], boundaries[P+1]]) on the device and, copied
   * after each run, in pinned host memory (the descriptor's segmentBoundaries points into it) */
  int *d_exps, *d_emi, *d_nwords, *d_vlc;
  double *d_pcs;
  long *d_upre; /* exclusive prefix of the chunks' work units (flat compress) */
  int *h_vlc;
  int vlc_valid;
  hipStream_t last_stream; /* stream of the last compute call */
  /* segment groups of the register FFT path: pass 1 of group g + 1 runs beside pass 2 of group
   * g on a second stream, so that pass 2 reads pass 1's output while it is still in the
   * last-level cache (fft_groups = 1: both passes over all segments, one launch each) */
  int fft_groups;
  hipStream_t stream2;
  hipEvent_t gev[SOI_MAX_FFT_GROUPS + 2];
  /* timing: NSTAGE+1 events per call, in a ring of NTSLOT calls */
  int timing;
  long ncalls;          /* calls recorded since timing was enabled */
  int filter_only_last;
  hipEvent_t ev[NTSLOT][NSTAGE + 1];
} soi_plan_t;

static int is_smooth(long n)
{
  if (n <= 0) return 0;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  while (n % 5 == 0) n /= 5;
  return n == 1;
}

void set_default_soi_descriptor(soi_desc_t *desc)
{
  clear_err();
  memset(desc, 0, sizeof(*desc));
  desc->comm_slot_ = 0;
  desc->P = 1;
  desc->rank = 0;
  desc->k = 1;
  desc->N = 0;
  desc->n_mu = 5;
  desc->d_mu = 4;
  desc->tau = 928 / 1024.;
  desc->sigma = 373.7314;
  desc->B = 72;
  desc->use_vlc = 0;
  desc->comm_to_comp_cost_ratio = 1;
  desc->alpha_tilde = NULL;
  desc->beta_tilde = NULL;
  desc->gamma_tilde = NULL;
  desc->precision = SOI_PREC_FP64;
  desc->window_mode = SOI_WINDOW_AUTO;
}

void soi_set_default_descriptor_precision(soi_desc_t *desc, int precision)
{
  set_default_soi_descriptor(desc);
  desc->precision = precision;
}

/* Filter launch configuration (see soi_filter.hip).  Window evaluation: exact windows with real
 * taps for (5,4) and B >= 60, where they agree with the reference's window values to ~1e-14 in
 * the output; otherwise the reference's own window values as complex taps. */
static int plan_filter(soi_plan_t *pl, int window_mode)
{
  SoiFilterArgs *a = &pl->fa;
  memset(a, 0, sizeof(*a));
  a->S = (int)pl->S;
  a->n_mu = pl->n_mu;
  a->d_mu = pl->d_mu;
  int rt_ok = (pl->n_mu == 5 && pl->d_mu == 4 && pl->B >= 60);
  if (window_mode == SOI_WINDOW_AUTO) {
    a->ct = !rt_ok;
    pl->ref_windows = !rt_ok;
  } else if (window_mode == SOI_WINDOW_REFERENCE) {
    a->ct = 1;
    pl->ref_windows = 1;
  } else if (window_mode == SOI_WINDOW_EXACT) {
    a->ct = !rt_ok;
    pl->ref_windows = 0;
  } else {
    set_err(SOI_ERR_ARGS, "unknown window_mode %d", window_mode);
    return -1;
  }
  if (soik_filter_shape(a->S, pl->n_mu, pl->d_mu, pl->B, a->ct, &a->NP, &a->BP, &a->NT, &a->U,
                        &a->W)) {
    set_err(SOI_ERR_ARGS, "no filter kernel for S=%ld, n_mu/d_mu=%d/%d, B=%d (S must be 4, 8, "
            "16, 32 or 64)", pl->S, pl->n_mu, pl->d_mu, pl->B);
    return -1;
  }
  a->M = pl->M;
  a->Mh = pl->Mh;
  a->nJ = pl->M / pl->d_mu;
  size_t lds = soik_filter_lds_bytes(a);
  if (lds > 160 * 1024) {
    set_err(SOI_ERR_ARGS, "filter tile needs %zu bytes of LDS (> 160 KiB) at S=%ld", lds, pl->S);
    return -1;
  }
  return 0;
}

/* largest T in {8,4,2,1} with n*T <= cap and T | other */
static int pick_tile(int n, long other, int cap)
{
  static const int ts[] = {8, 4, 2, 1};
  for (int t = 0; t < 4; ++t)
    if ((long)n * ts[t] <= cap && other % ts[t] == 0) return ts[t];
  return 0;
}

/* Register-FFT alternatives that exist for (n1, n2): column tile 8 or 4 with its kernel
 * variant, row tile 8 or 4.  The planned choice is entry 0 of the table. */
static void fft_variants(soi_plan_t *pl)
{
  int cols[6][2], nc = 0, rows[2], nr = 0;
  const int n1 = pl->n1, n2 = pl->n2;
  const int c4 = n1 == 1024 || n1 == 1280 || n1 == 2048 || n1 == 2560;
  cols[nc][0] = pl->Tc, cols[nc][1] = pl->clay, ++nc;
  if (soik_fft_cols_short_tile(n1)) { /* short columns: one configuration */
    pl->nvar = 1;
    pl->vtab[0][0] = pl->Tc, pl->vtab[0][1] = pl->clay, pl->vtab[0][2] = pl->Tr;
    pl->var = 0;
    return;
  }
  if (n2 % 8 == 0 && !(pl->Tc == 8 && pl->clay == 0)) cols[nc][0] = 8, cols[nc][1] = 0, ++nc;
  if (c4 && n2 % 4 == 0 && !(pl->Tc == 4 && pl->clay == 0)) cols[nc][0] = 4, cols[nc][1] = 0, ++nc;
  if ((n1 == 2048 || n1 == 2560) && n2 % 8 == 0 && pl->clay != 1)
    cols[nc][0] = 8, cols[nc][1] = 1, ++nc;
  if (n1 == 2560 && n2 % 4 == 0 && pl->clay != 2 && nc < 5) cols[nc][0] = 4, cols[nc][1] = 2, ++nc;
  if (n1 == 2560 && pl->f32 && n2 % 8 == 0 && pl->clay != 3 && nc < 6)
    cols[nc][0] = 8, cols[nc][1] = 3, ++nc;
  if (n1 == 2560 && !pl->f32 && n2 % 4 == 0 && pl->clay != 4 && nc < 6)
    cols[nc][0] = 4, cols[nc][1] = 4, ++nc;
  rows[nr++] = pl->Tr;
  if (pl->Tr != 8 && n2 != 2560 && n1 % 8 == 0) rows[nr++] = 8;
  if (pl->Tr != 4 && (n2 == 1024 || n2 == 1280 || n2 == 2048 || n2 == 2560) && n1 % 4 == 0)
    rows[nr++] = 4;
  pl->nvar = 0;
  for (int r = 0; r < nr; ++r)
    for (int c = 0; c < nc; ++c) {
      int *v = pl->vtab[pl->nvar++];
      v[0] = cols[c][0], v[1] = cols[c][1], v[2] = rows[r];
    }
  pl->var = 0;
}

static int plan_fft(soi_plan_t *pl)
{
  const long Mh = pl->Mh;
  const int CAP = 5120; /* complex values per block: 80 KiB of LDS, two blocks per CU */
  if (!is_smooth(Mh)) {
    set_err(SOI_ERR_ARGS, "M_hat=%ld is not of the form 2^a 3^b 5^c", Mh);
    return -1;
  }
  if (Mh + 1 <= CAP) {
    pl->nvar = 1;
    pl->var = 0;
    pl->fft_single = 1;
    pl->single_count = (int)(CAP / (Mh + 1));
    if (pl->single_count > pl->S) pl->single_count = (int)pl->S;
    if (pl->single_count < 1) pl->single_count = 1;
    return 0;
  }
  pl->fft_single = 0;
  /* register kernels: columns of n1 and rows of n2 both have one, 8-wide tiles on both sides */
  {
    long best = LONG_MAX;
    int bn1 = 0;
    /* SOI_FFT_N1 forces the column length (tuning) */
    const char *en = getenv("SOI_FFT_N1");
    long force = en ? atol(en) : 0;
    for (long n1 = 2; n1 <= 4096; ++n1) {
      if (Mh % n1) continue;
      long n2 = Mh / n1;
      const int st = soik_fft_cols_short_tile((int)n1);
      if (n2 % (st ? st : 8) || n2 > 4096) continue;
      if (!soik_fft_reg_supported((int)n1, 0) || !soik_fft_reg_supported((int)n2, 1)) continue;
      if (st && n1 % 8 && n2 != 256) continue; /* 10-wide rows exist for 256 points only */
      long cost = (n1 > n2 ? n1 : n2);
      if (force == n1) cost = 0;
      if (cost <= best) {   /* ties: the longer column */
        best = cost;
        bn1 = (int)n1;
      }
    }
    if (bn1) {
      pl->fft_reg = 1;
      pl->n1 = bn1;
      pl->n2 = (int)(Mh / bn1);
      /* tile widths: 8 by default; SOI_FFT_TC / SOI_FFT_TR (4 or 8) override for tuning */
      pl->Tc = pl->Tr = 8;
      const char *ec = getenv("SOI_FFT_TC"), *er = getenv("SOI_FFT_TR");
      if (ec && (atoi(ec) == 4 || atoi(ec) == 8)) pl->Tc = atoi(ec);
      if (er && (atoi(er) == 4 || atoi(er) == 8)) pl->Tr = atoi(er);
      if (pl->n2 % pl->Tc || pl->n1 % pl->Tr) pl->Tc = pl->Tr = 8;
      if (pl->n2 == 2560) pl->Tr = 4; /* 2560-point rows exist with 4-wide tiles only */
      const int stile = soik_fft_cols_short_tile(pl->n1);
      if (stile) { /* short columns: their own tile width; rows 10-wide unless n1 % 8 == 0 */
        pl->Tc = stile;
        pl->Tr = pl->n1 % 8 ? 10 : 8;
      }
      /* column-pass kernel variant: 1 = one padded LDS run per sequence (8-wide tiles of 2048
       * or 2560); 2 = one 4-wide block per CU with the whole register budget (2560: 20
       * values per thread do not fit 1024-thread blocks); 4 = the same exchanging whole
       * complex128 values in all 160 KiB of LDS (2560, the complex128 default: one round per
       * exchange instead of a real and an imaginary one, 1.59 -> 1.51 ms at N = 2^27).
       * Complex64 2560-point columns default to 4-wide tiles in two exchange groups, two
       * blocks per CU (their whole-complex exchanges fit in 44 KB).  SOI_FFT_CVAR overrides
       * for tuning */
      const char *el = getenv("SOI_FFT_CVAR");
      int cv = el && !stile ? atoi(el) : (pl->n1 == 2560 && !pl->f32 ? 4 : 0);
      if (!el && !ec && pl->n1 == 2560 && pl->f32 && pl->n2 % 4 == 0) pl->Tc = 4;
      pl->clay = 0;
      if (cv == 1 && pl->Tc == 8 && (pl->n1 == 2048 || pl->n1 == 2560)) pl->clay = 1;
      if (cv == 2 && pl->n1 == 2560) {
        pl->clay = 2;
        pl->Tc = 4;
      }
      if (cv == 3 && pl->n1 == 2560 && pl->f32) {
        pl->clay = 3;
        pl->Tc = 8;
      }
      if (cv == 4 && pl->n1 == 2560 && !pl->f32) {
        pl->clay = 4;
        pl->Tc = 4;
      }
      int lb = 0;
      while ((1L << lb) * (1L << lb) < Mh) ++lb;
      pl->B1 = 1 << lb;
      pl->lb1 = lb;
      /* tiled pass-1 output: the default for complex64 (measured faster there; for
       * complex128 the tiled row pass spills and the split stores are slower), SOI_FFT_TILED
       * overrides */
      const char *et = getenv("SOI_FFT_TILED");
      pl->tiled = (et ? atoi(et) != 0 : pl->f32) && pl->n1 % 8 == 0 &&
                  soik_fft_rows_tiled_supported(pl->n2, pl->f32);
      fft_variants(pl);
      return 0;
    }
  }
  /* generic LDS kernels: n1 x n2 = Mh, both smooth, column tile Tc >= 2 and row tile Tr >= 2
   * preferred; minimise max(n1, n2) with n1 a power of two when possible */
  long best_cost = LONG_MAX;
  int bn1 = 0, bn2 = 0, bTc = 0, bTr = 0;
  for (long n1 = 2; n1 <= CAP; ++n1) {
    if (Mh % n1) continue;
    long n2 = Mh / n1;
    if (n2 > CAP || !is_smooth(n1) || !is_smooth(n2)) continue;
    int Tc = pick_tile((int)n1, n2, CAP), Tr = pick_tile((int)n2, n1, CAP);
    if (!Tc || !Tr) continue;
    long cost = (n1 > n2 ? n1 : n2) * 4 + (Tc < 2) * 100000 + (Tr < 2) * 100000 +
                ((n1 & (n1 - 1)) ? 1 : 0);
    if (cost < best_cost) {
      best_cost = cost;
      bn1 = (int)n1;
      bn2 = (int)n2;
      bTc = Tc;
      bTr = Tr;
    }
  }
  if (!bn1) {
    set_err(SOI_ERR_ARGS, "M_hat=%ld is too long for the two-pass segment FFT (max %d^2)", Mh,
            CAP);
    return -1;
  }
  pl->n1 = bn1;
  pl->n2 = bn2;
  pl->Tc = bTc;
  pl->Tr = bTr;
  pl->nvar = 1;
  pl->var = 0;
  pl->vtab[0][0] = bTc, pl->vtab[0][1] = 0, pl->vtab[0][2] = bTr;
  int b1 = 1, lb = 0;
  while ((long)b1 * b1 < Mh) b1 <<= 1, ++lb;
  pl->B1 = b1;
  pl->lb1 = lb;
  return 0;
}

static void free_plan(soi_plan_t *pl)
{
  if (!pl) return;
  if (pl->device >= 0) (void)hipSetDevice(pl->device);
  (void)hipFree(pl->taps);
  (void)hipFree(pl->post);
  (void)hipFree(pl->twS);
  (void)hipFree(pl->taps32);
  (void)hipFree(pl->wt);
  (void)hipFree(pl->post32);
  (void)hipFree(pl->twS32);
  (void)hipFree(pl->tw1_32);
  (void)hipFree(pl->tw2_32);
  (void)hipFree(pl->thi32);
  (void)hipFree(pl->tlo32);
  (void)hipFree(pl->tw1);
  (void)hipFree(pl->tw2);
  (void)hipFree(pl->thi);
  (void)hipFree(pl->tlo);
  (void)hipFree(pl->io);
  (void)hipFree(pl->tbuf);
  if (pl->owns_at) (void)hipFree(pl->at);
  (void)hipFree(pl->wtab);
  (void)hipFree(pl->winv);
  (void)hipFree(pl->delta);
  (void)hipFree(pl->d_exps);
  (void)hipFree(pl->d_emi);
  (void)hipFree(pl->d_nwords);
  (void)hipFree(pl->d_vlc);
  (void)hipFree(pl->d_pcs);
  (void)hipFree(pl->d_upre);
  if (pl->h_vlc) (void)hipHostFree(pl->h_vlc);
  for (int c = 0; c < NTSLOT; ++c)
    for (int e = 0; e <= NSTAGE; ++e)
      if (pl->ev[c][e]) (void)hipEventDestroy(pl->ev[c][e]);
  for (int e = 0; e < SOI_MAX_FFT_GROUPS + 2; ++e)
    if (pl->gev[e]) (void)hipEventDestroy(pl->gev[e]);
  if (pl->stream2) (void)hipStreamDestroy(pl->stream2);
  if (pl->stream) (void)hipStreamDestroy(pl->stream);
  free(pl);
}

void init_soi_descriptor(soi_desc_t *d, MPI_Comm comm, cfft_size_t k)
{
  clear_err();
  soi_plan_t *pl = NULL;
  d->comm = comm;
  d->k = k;
  if (d->plan) {
    set_err(SOI_ERR_STATE, "descriptor already initialised; call free_soi_descriptor first");
    return;
  }
  if (d->P < 1 || k < 1 || k > 1024 || d->N == 0) {
    set_err(SOI_ERR_ARGS, "invalid N=%zu, P=%d, k=%zu (N>0, P>=1, 1<=k<=1024)", d->N, d->P, k);
    return;
  }
  if (d->rank != 0) {
    set_err(SOI_ERR_ARGS, "rank must be 0: one GPU runs all P ranks as virtual ranks");
    return;
  }
  if (!((d->n_mu == 5 && d->d_mu == 4) || (d->n_mu == 8 && d->d_mu == 7))) {
    set_err(SOI_ERR_UNSUPPORTED_MU,
            "Unsupported n_mu and d_mu (%d/%d). Try n_mu=5 && d_mu=4 or n_mu=8 && d_mu=7",
            d->n_mu, d->d_mu);
    return;
  }
  if (d->precision != SOI_PREC_FP64 && d->precision != SOI_PREC_FP32) {
    set_err(SOI_ERR_ARGS, "unknown precision %d", d->precision);
    return;
  }
  if (d->precision == SOI_PREC_FP32 && d->use_vlc) {
    set_err(SOI_ERR_ARGS, "use_vlc needs precision SOI_PREC_FP64: the VLC words encode FP64 "
            "values (compress.c)");
    return;
  }
  long S = (long)k * d->P;
  if (d->N % S) {
    set_err(SOI_ERR_ARGS, "S=k*P=%ld must divide N=%zu", S, d->N);
    return;
  }
  long M = (long)d->N / S;
  if (S % 4 || !is_smooth(S)) {
    set_err(SOI_ERR_ARGS, "S=%ld must be a multiple of 4 (pfft.c:192) of the form 2^a 3^b 5^c", S);
    return;
  }
  if (M % ((long)d->P * d->d_mu)) {
    set_err(SOI_ERR_ARGS, "P*d_mu=%ld must divide M=N/S=%ld", (long)d->P * d->d_mu, M);
    return;
  }
  if (M / d->P < (long)d->B) {
    set_err(SOI_ERR_TOO_SMALL, "input size too small: M/P=%ld < B=%zu", M / d->P, d->B);
    return;
  }
  if (d->B < 2 || d->B <= (cfft_size_t)d->d_mu || d->B > 80) {
    set_err(SOI_ERR_ARGS, "B=%zu must exceed d_mu=%d and be at most 80", d->B, d->d_mu);
    return;
  }

  pl = (soi_plan_t *)calloc(1, sizeof(soi_plan_t));
  if (!pl) {
    set_err(SOI_ERR_ALLOC, "host allocation failed");
    return;
  }
  pl->device = -1;
  HIPCHK(hipGetDevice(&pl->device));
  pl->N = (long)d->N;
  pl->P = d->P;
  pl->k = (int)k;
  pl->S = S;
  pl->M = M;
  pl->n_mu = d->n_mu;
  pl->d_mu = d->d_mu;
  pl->B = (int)d->B;
  pl->Mh = (long)d->n_mu * M / d->d_mu;
  pl->l = pl->Mh / d->P;
  pl->f32 = d->precision == SOI_PREC_FP32;
  if (plan_filter(pl, d->window_mode) || plan_fft(pl)) goto fail_quiet;
  HIPCHK(hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking));
  {
    /* SOI_FFT_GROUPS: segment groups of the two-stream FFT schedule (a divisor of S) */
    const char *eg = getenv("SOI_FFT_GROUPS");
    int ng = eg ? atoi(eg) : 1;
    if (ng < 1 || ng > SOI_MAX_FFT_GROUPS || S % ng || !pl->fft_reg) ng = 1;
    pl->fft_groups = ng;
    if (ng > 1) {
      HIPCHK(hipStreamCreateWithFlags(&pl->stream2, hipStreamNonBlocking));
      for (int e = 0; e < ng + 2; ++e)
        HIPCHK(hipEventCreateWithFlags(&pl->gev[e], hipEventDisableTiming));
    }
  }

  /* segment boundaries p*k (pfft.c:144-147) */
  d->segmentBoundaries = (int *)malloc(sizeof(int) * (d->P + 1));
  for (int p = 0; p <= d->P; ++p) d->segmentBoundaries[p] = p * (int)k;

  {
    int B_pad = pl->fa.NP * pl->fa.BP;
    size_t ntap = (size_t)d->n_mu * S * B_pad;
    HIPCHK(hipMalloc((void **)&pl->taps, ntap * (pl->fa.ct ? 2 : 1) * sizeof(double)));
    HIPCHK(hipMalloc((void **)&pl->post, (size_t)d->n_mu * S * sizeof(soi_c64_t)));
    HIPCHK(hipMalloc((void **)&pl->wtab, (size_t)d->n_mu * d->B * S * sizeof(soi_c64_t)));
    HIPCHK(hipMalloc((void **)&pl->winv, (size_t)M * sizeof(soi_c64_t)));
    HIPCHK(hipMalloc((void **)&pl->twS, (size_t)S * sizeof(soi_c64_t)));
    d->w_dup = pl->taps;
    const size_t esz = pl->f32 ? sizeof(soi_c32_t) : sizeof(soi_c64_t);
    const size_t at_bytes = (size_t)S * pl->Mh * esz;
    if (d->alpha_tilde) {
      /* a preallocated alpha_tilde holds all S segments ([S][M_hat], S*M_hat values): the
       * reference's per-rank M_hat*k (pfft.c:161-163) is P times too small here */
      void *base = NULL;
      size_t size = 0;
      if (hipMemGetAddressRange(&base, &size, d->alpha_tilde) != hipSuccess || !base) {
        (void)hipGetLastError();
        set_err(SOI_ERR_ARGS, "preallocated alpha_tilde is not device memory");
        goto fail_quiet;
      }
      const size_t avail = (size_t)((char *)base + size - (char *)d->alpha_tilde);
      if (avail < at_bytes) {
        set_err(SOI_ERR_ARGS, "preallocated alpha_tilde holds %zu bytes; S*M_hat values need "
                "%zu (S=%ld, M_hat=%ld)", avail, at_bytes, S, pl->Mh);
        goto fail_quiet;
      }
      pl->at = d->alpha_tilde;
    } else {
      hipError_t e = hipMalloc(&pl->at, at_bytes);
      if (e != hipSuccess) {
        pl->at = NULL;
        set_err(SOI_ERR_ALLOC, "Failed to allocate d->alpha_tilde (%zu bytes): %s", at_bytes,
                hipGetErrorString(e));
        goto fail_quiet;
      }
      pl->owns_at = 1;
    }
    if (pl->tiled) {
      hipError_t e = hipMalloc(&pl->tbuf, (size_t)S * pl->Mh * esz);
      if (e != hipSuccess) {
        pl->tbuf = NULL;
        set_err(SOI_ERR_ALLOC, "Failed to allocate the FFT tile buffer (%zu bytes): %s",
                (size_t)S * pl->Mh * esz, hipGetErrorString(e));
        goto fail_quiet;
      }
    }
    KCHK(soik_make_tables(pl->taps, pl->post, pl->wtab, pl->winv, (int)S,
                          d->n_mu, d->d_mu, (int)d->B, B_pad, d->tau, d->sigma, pl->N, M, pl->Mh,
                          pl->ref_windows, pl->fa.ct, pl->stream));
    KCHK(soik_make_twiddles(pl->twS, S, S, 1, pl->stream));
    pl->fa.taps = pl->taps;
    pl->fa.post = pl->post;
    pl->fa.twS = pl->twS;
    if (pl->fft_single) {
      HIPCHK(hipMalloc((void **)&pl->tw1, (size_t)pl->Mh * sizeof(soi_c64_t)));
      KCHK(soik_make_twiddles(pl->tw1, pl->Mh, pl->Mh, 1, pl->stream));
    } else {
      long nhi = (pl->Mh + pl->B1 - 1) / pl->B1;
      HIPCHK(hipMalloc((void **)&pl->tw1, (size_t)pl->n1 * sizeof(soi_c64_t)));
      HIPCHK(hipMalloc((void **)&pl->tw2, (size_t)pl->n2 * sizeof(soi_c64_t)));
      HIPCHK(hipMalloc((void **)&pl->thi, (size_t)nhi * sizeof(soi_c64_t)));
      HIPCHK(hipMalloc((void **)&pl->tlo, (size_t)pl->B1 * sizeof(soi_c64_t)));
      KCHK(soik_make_twiddles(pl->tw1, pl->n1, pl->n1, 1, pl->stream));
      KCHK(soik_make_twiddles(pl->tw2, pl->n2, pl->n2, 1, pl->stream));
      KCHK(soik_make_twiddles(pl->thi, nhi, pl->Mh, pl->B1, pl->stream));
      KCHK(soik_make_twiddles(pl->tlo, pl->B1, pl->Mh, 1, pl->stream));
    }
    if (!pl->f32) {
      /* The complex128 filter contraction on MFMA (v_mfma_f64_16x16x4_f64) where the shape
       * allows: the default.  SOI_FILTER_MFMA64=0 selects the VALU form. */
      const char *em = getenv("SOI_FILTER_MFMA64");
      if ((!em || atoi(em) != 0) &&
          soik_filter_mfma_supported((int)S, d->n_mu, d->d_mu, pl->fa.ct, B_pad, 3)) {
        HIPCHK(hipMalloc(&pl->wt, soik_filter_mfma_table_bytes((int)S, B_pad, 3)));
        KCHK(soik_filter_mfma_table((const double *)pl->taps, pl->wt, (int)S, B_pad, 3,
                                    pl->stream));
        pl->fa.wt = pl->wt;
        pl->fa.mfma = 3;
      }
    }
    if (pl->f32) {
      /* FP32 copies of the tables, rounded from the FP64 values generated above */
      size_t nt = ntap * (pl->fa.ct ? 2 : 1); /* doubles; even, since S is a multiple of 4 */
      HIPCHK(hipMalloc((void **)&pl->taps32, nt * sizeof(float)));
      KCHK(soik_f64_to_f32((const soi_c64_t *)pl->taps, pl->taps32, (long)(nt / 2), pl->stream));
      HIPCHK(hipMalloc((void **)&pl->post32, (size_t)d->n_mu * S * sizeof(soi_c32_t)));
      KCHK(soik_f64_to_f32(pl->post, pl->post32, (long)d->n_mu * S, pl->stream));
      HIPCHK(hipMalloc((void **)&pl->twS32, (size_t)S * sizeof(soi_c32_t)));
      KCHK(soik_f64_to_f32(pl->twS, pl->twS32, S, pl->stream));
      pl->fa.taps = pl->taps32;
      pl->fa.post = pl->post32;
      pl->fa.twS = pl->twS32;
      pl->fa.f32 = 1;
      /* The filter contraction on MFMA with FP32 operands (v_mfma_f32_16x16x4_f32) where the
       * shape allows: the default, faster than the FP32 VALU form.  SOI_FILTER_MFMA=0 selects
       * the VALU form, =1 the split-bf16 MFMA form (slower than both). */
      const char *em = getenv("SOI_FILTER_MFMA");
      const int mode = em ? atoi(em) : 2;
      if ((mode == 1 || mode == 2) &&
          soik_filter_mfma_supported((int)S, d->n_mu, d->d_mu, pl->fa.ct, B_pad, mode)) {
        HIPCHK(hipMalloc(&pl->wt, soik_filter_mfma_table_bytes((int)S, B_pad, mode)));
        KCHK(soik_filter_mfma_table((const double *)pl->taps, pl->wt, (int)S, B_pad, mode,
                                    pl->stream));
        pl->fa.wt = pl->wt;
        pl->fa.mfma = mode;
      }
      if (pl->fft_single) {
        HIPCHK(hipMalloc((void **)&pl->tw1_32, (size_t)pl->Mh * sizeof(soi_c32_t)));
        KCHK(soik_f64_to_f32(pl->tw1, pl->tw1_32, pl->Mh, pl->stream));
      } else {
        long nhi = (pl->Mh + pl->B1 - 1) / pl->B1;
        HIPCHK(hipMalloc((void **)&pl->tw1_32, (size_t)pl->n1 * sizeof(soi_c32_t)));
        HIPCHK(hipMalloc((void **)&pl->tw2_32, (size_t)pl->n2 * sizeof(soi_c32_t)));
        HIPCHK(hipMalloc((void **)&pl->thi32, (size_t)nhi * sizeof(soi_c32_t)));
        HIPCHK(hipMalloc((void **)&pl->tlo32, (size_t)pl->B1 * sizeof(soi_c32_t)));
        KCHK(soik_f64_to_f32(pl->tw1, pl->tw1_32, pl->n1, pl->stream));
        KCHK(soik_f64_to_f32(pl->tw2, pl->tw2_32, pl->n2, pl->stream));
        KCHK(soik_f64_to_f32(pl->thi, pl->thi32, nhi, pl->stream));
        KCHK(soik_f64_to_f32(pl->tlo, pl->tlo32, pl->B1, pl->stream));
      }
    }
    if (d->use_vlc) {
      long nch = S * d->P;
      size_t nvlc = (size_t)(1 + 2 * S + d->P + 1);
      HIPCHK(hipMalloc((void **)&pl->delta, (size_t)nch * 4 * pl->l * sizeof(int) + 256));
      HIPCHK(hipMalloc((void **)&pl->d_exps, nch * sizeof(int)));
      HIPCHK(hipMalloc((void **)&pl->d_emi, nch * sizeof(int)));
      HIPCHK(hipMalloc((void **)&pl->d_nwords, nch * sizeof(int)));
      HIPCHK(hipMalloc((void **)&pl->d_vlc, nvlc * sizeof(int)));
      HIPCHK(hipMalloc((void **)&pl->d_pcs, (size_t)(S + 1) * sizeof(double)));
      HIPCHK(hipMalloc((void **)&pl->d_upre, (size_t)(nch + 1) * sizeof(long)));
      HIPCHK(hipHostMalloc((void **)&pl->h_vlc, nvlc * sizeof(int), hipHostMallocDefault));
      /* segmentBoundaries lives in the pinned plan copy, so that the run's device plan reaches
       * it by an asynchronous copy */
      int *sb = pl->h_vlc + 1 + 2 * S;
      memcpy(sb, d->segmentBoundaries, sizeof(int) * (d->P + 1));
      free(d->segmentBoundaries);
      d->segmentBoundaries = sb;
    }
    HIPCHK(hipStreamSynchronize(pl->stream));
  }
  d->alpha_tilde = (cfft_complex_t *)pl->at;
  d->w = (cfft_complex_t *)pl->wtab;
  d->W_inv = (cfft_complex_t *)pl->winv;
  d->delta = pl->delta;
  d->plan = pl;
  return;
fail:
fail_quiet:
  if (pl) {
    d->w_dup = NULL;
    if (!pl->h_vlc || d->segmentBoundaries != pl->h_vlc + 1 + 2 * pl->S) free(d->segmentBoundaries);
    d->segmentBoundaries = NULL;
    free_plan(pl);
  }
}

void free_soi_descriptor(soi_desc_t *d)
{
  clear_err();
  soi_plan_t *pl = (soi_plan_t *)d->plan;
  if (!pl) {
    set_err(SOI_ERR_STATE, "descriptor is not initialised");
    return;
  }
  (void)hipSetDevice(pl->device);
  (void)hipStreamSynchronize(pl->stream);
  d->alpha_tilde = NULL;
  d->beta_tilde = NULL;
  d->gamma_tilde = NULL;
  d->w = NULL;
  d->W_inv = NULL;
  d->w_dup = NULL;
  d->delta = NULL;
  d->epsilon = NULL;
  if (!pl->h_vlc) free(d->segmentBoundaries); /* else it points into the plan's pinned copy */
  d->segmentBoundaries = NULL;
  free_plan(pl);
  d->plan = NULL;
}

/* device pointer?  (hipPointerGetAttributes fails or says host for plain host memory) */
static int is_device_ptr(const void *p)
{
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

static void mark(soi_plan_t *pl, int e, hipStream_t st)
{
  if (pl->timing) (void)hipEventRecord(pl->ev[pl->ncalls % NTSLOT][e], st);
}

static int run_filter(soi_desc_t *d, soi_plan_t *pl, const void *x, SoiBatch b, hipStream_t st)
{
  pl->fa.x = x;
  pl->fa.at = pl->at;
  pl->fa.b = b;
  int e = soik_filter_sdft(&pl->fa, st);
  pl->fa.b = (SoiBatch){0};
  if (e) {
    set_err(SOI_ERR_HIP, "filter kernel launch failed: %s", hipGetErrorString((hipError_t)e));
    return -1;
  }
  return 0;
}

/* The VLC branch (pfft.c:316-353, 390-440, 476-543) without leaving the device: per-chunk
 * maximum exponents, then the plan (segment and global maxima, send counts, segment
 * boundaries) in k_vlc_plan, compress and decompress reading the global maximum from device
 * memory, and an asynchronous copy of the plan into pinned host memory (segmentBoundaries,
 * soi_vlc_exponents). */
static int run_vlc(soi_desc_t *d, soi_plan_t *pl, hipStream_t st)
{
  const int S = (int)pl->S, P = pl->P;
  const long l = pl->l;
  const int nch = S * P;
  double *at = (double *)pl->at;
  /* chunk (s, p) = alpha_tilde[s][p*l .. (p+1)*l): 2l doubles at at + 2*(s*Mh + p*l);
   * with Mh = P*l the chunks are contiguous in (s, p) order, stride 2l doubles */
  KCHK(soik_max_exponent(at, (int)(2 * l), nch, pl->d_exps, st));
  KCHK(soik_vlc_plan(pl->d_exps, S, P, pl->k, l, pl->Mh, d->comm_to_comp_cost_ratio, pl->d_emi,
                     pl->d_vlc, pl->d_pcs, pl->d_upre, pl->d_nwords, st));
  KCHK(soik_compress_flat(pl->delta, 4 * l, at, (int)(2 * l), nch, pl->d_vlc, pl->d_emi,
                          pl->d_upre, st));
  KCHK(soik_decompress_flat(at, pl->delta, 4 * l, (int)(2 * l), nch, pl->d_vlc, pl->d_emi, st));
  HIPCHK(hipMemcpyAsync(pl->h_vlc, pl->d_vlc, sizeof(int) * (size_t)(1 + 2 * S + P + 1),
                        hipMemcpyDeviceToHost, st));
  pl->vlc_valid = 1;
  return 0;
fail:
  return -1;
}

static int run_fft_f32(soi_desc_t *d, soi_plan_t *pl, soi_c32_t *X, SoiBatch b, hipStream_t st)
{
  soi_c32_t *at = (soi_c32_t *)pl->at;
  const soi_c64_t *winv = pl->winv;
  if (pl->fft_single) {
    mark(pl, 2, st);
    KCHK(soik_fft_single_f32(at, X, pl->tw1_32, winv, pl->Mh, pl->M, (int)pl->S,
                             pl->single_count, b, st));
    mark(pl, 3, st);
  } else if (pl->fft_reg) {
    soi_c32_t *gt = pl->tiled ? (soi_c32_t *)pl->tbuf : NULL;
    const int ng = b.n > 0 ? 1 : pl->fft_groups, G = (int)pl->S / ng;
    hipStream_t s2 = ng > 1 ? pl->stream2 : st;
    if (ng > 1) {
      HIPCHK(hipEventRecord(pl->gev[0], st));
      HIPCHK(hipStreamWaitEvent(s2, pl->gev[0], 0));
    }
    for (int g = 0; g < ng; ++g) {
      const long so = (long)g * G;
      KCHK(soik_fft_cols_reg_f32(at + so * pl->Mh, pl->tw1_32, pl->thi32, pl->tlo32, pl->n1,
                                 pl->n2, pl->lb1, pl->Mh, G, pl->Tc, pl->clay,
                                 gt ? gt + so * pl->Mh : NULL, b, st));
      if (ng > 1) {
        HIPCHK(hipEventRecord(pl->gev[1 + g], st));
        HIPCHK(hipStreamWaitEvent(s2, pl->gev[1 + g], 0));
      } else {
        mark(pl, 3, st);
      }
      if (gt)
        KCHK(soik_fft_rows_tiled_f32(gt + so * pl->Mh, X + so * pl->M, pl->tw2_32, winv, pl->n1,
                                     pl->n2, pl->Mh, pl->M, G, b, s2));
      else
        KCHK(soik_fft_rows_reg_f32(at + so * pl->Mh, X + so * pl->M, pl->tw2_32, winv, pl->n1,
                                   pl->n2, pl->Mh, pl->M, G, pl->Tr, b, s2));
    }
    if (ng > 1) {
      HIPCHK(hipEventRecord(pl->gev[ng + 1], s2));
      HIPCHK(hipStreamWaitEvent(st, pl->gev[ng + 1], 0));
      mark(pl, 3, st);
    }
  } else {
    KCHK(soik_fft_col_f32(at, pl->tw1_32, pl->thi32, pl->tlo32, pl->n1, pl->n2, pl->Tc, pl->B1,
                          pl->Mh, (int)pl->S, b, st));
    mark(pl, 3, st);
    KCHK(soik_fft_row_f32(at, X, pl->tw2_32, winv, pl->n1, pl->n2, pl->Tr, pl->Mh, pl->M,
                          (int)pl->S, b, st));
  }
  return 0;
fail:
  return -1;
}

static int run_fft(soi_desc_t *d, soi_plan_t *pl, void *Xv, SoiBatch b, hipStream_t st)
{
  if (pl->f32) return run_fft_f32(d, pl, (soi_c32_t *)Xv, b, st);
  soi_c64_t *X = (soi_c64_t *)Xv;
  soi_c64_t *at = (soi_c64_t *)pl->at;
  if (pl->fft_single) {
    mark(pl, 2, st);
    KCHK(soik_fft_single(at, X, pl->tw1, pl->winv, pl->Mh, pl->M, (int)pl->S,
                         pl->single_count, b, st));
    mark(pl, 3, st);
  } else if (pl->fft_reg) {
    soi_c64_t *gt = pl->tiled ? (soi_c64_t *)pl->tbuf : NULL;
    const int ng = b.n > 0 ? 1 : pl->fft_groups, G = (int)pl->S / ng;
    hipStream_t s2 = ng > 1 ? pl->stream2 : st;
    if (ng > 1) {
      HIPCHK(hipEventRecord(pl->gev[0], st));
      HIPCHK(hipStreamWaitEvent(s2, pl->gev[0], 0));
    }
    for (int g = 0; g < ng; ++g) {
      const long so = (long)g * G;
      KCHK(soik_fft_cols_reg(at + so * pl->Mh, pl->tw1, pl->thi, pl->tlo, pl->n1, pl->n2,
                             pl->lb1, pl->Mh, G, pl->Tc, pl->clay, gt ? gt + so * pl->Mh : NULL,
                             b, st));
      if (ng > 1) {
        HIPCHK(hipEventRecord(pl->gev[1 + g], st));
        HIPCHK(hipStreamWaitEvent(s2, pl->gev[1 + g], 0));
      } else {
        mark(pl, 3, st);
      }
      if (gt)
        KCHK(soik_fft_rows_tiled(gt + so * pl->Mh, X + so * pl->M, pl->tw2, pl->winv, pl->n1,
                                 pl->n2, pl->Mh, pl->M, G, b, s2));
      else
        KCHK(soik_fft_rows_reg(at + so * pl->Mh, X + so * pl->M, pl->tw2, pl->winv, pl->n1,
                               pl->n2, pl->Mh, pl->M, G, pl->Tr, b, s2));
    }
    if (ng > 1) {
      HIPCHK(hipEventRecord(pl->gev[ng + 1], s2));
      HIPCHK(hipStreamWaitEvent(st, pl->gev[ng + 1], 0));
      mark(pl, 3, st);
    }
  } else {
    KCHK(soik_fft_col(at, pl->tw1, pl->thi, pl->tlo, pl->n1, pl->n2, pl->Tc, pl->B1, pl->Mh,
                      (int)pl->S, b, st));
    mark(pl, 3, st);
    KCHK(soik_fft_row(at, X, pl->tw2, pl->winv, pl->n1, pl->n2, pl->Tr,
                      pl->Mh, pl->M, (int)pl->S, b, st));
  }
  return 0;
fail:
  return -1;
}

static void compute_impl(soi_desc_t *d, cfft_complex_t *alpha_dt, void *stream, int sync,
                         int filter_only)
{
  clear_err();
  soi_plan_t *pl = (soi_plan_t *)d->plan;
  if (!pl) {
    set_err(SOI_ERR_STATE, "descriptor is not initialised (call init_soi_descriptor)");
    return;
  }
  if (!alpha_dt) {
    set_err(SOI_ERR_ARGS, "alpha_dt is NULL");
    return;
  }
  hipStream_t st = stream ? (hipStream_t)stream : pl->stream;
  HIPCHK(hipSetDevice(pl->device));
  pl->last_stream = st;
  {
    int dev = is_device_ptr(alpha_dt);
    soi_c64_t *buf = (soi_c64_t *)alpha_dt;
    size_t bytes = (size_t)pl->N * (pl->f32 ? sizeof(soi_c32_t) : sizeof(soi_c64_t));
    if (!dev) {
      if (!pl->io) {
        hipError_t e = hipMalloc((void **)&pl->io, bytes);
        if (e != hipSuccess) {
          pl->io = NULL;
          set_err(SOI_ERR_ALLOC, "staging allocation of %zu bytes failed", bytes);
          return;
        }
      }
      HIPCHK(hipMemcpyAsync(pl->io, alpha_dt, bytes, hipMemcpyHostToDevice, st));
      buf = pl->io;
    }
    pl->vlc_valid = 0;
    mark(pl, 0, st);
    if (run_filter(d, pl, buf, (SoiBatch){0}, st)) return;
    mark(pl, 1, st);
    if (!filter_only) {
      if (d->use_vlc && run_vlc(d, pl, st)) return;
      if (!pl->fft_single) mark(pl, 2, st);
      if (run_fft(d, pl, buf, (SoiBatch){0}, st)) return;
      mark(pl, 4, st);
    }
    if (pl->timing) {
      pl->filter_only_last = filter_only;
      pl->ncalls++;
    }
    if (!dev) {
      if (!filter_only)
        HIPCHK(hipMemcpyAsync(alpha_dt, pl->io, bytes, hipMemcpyDeviceToHost, st));
      sync = 1;
    }
    if (sync) HIPCHK(hipStreamSynchronize(st));
  }
  return;
fail:
  return;
}

void compute_soi(soi_desc_t *d, cfft_complex_t *alpha_dt) { compute_impl(d, alpha_dt, NULL, 1, 0); }

void compute_soi_async(soi_desc_t *d, cfft_complex_t *alpha_dt, void *stream)
{
  compute_impl(d, alpha_dt, stream, 0, 0);
}

void parallel_filter_subsampling(soi_desc_t *desc, cfft_complex_t *alpha_dt)
{
  compute_impl(desc, alpha_dt, NULL, 1, 1);
}

void *soi_stream(soi_desc_t *d)
{
  soi_plan_t *pl = (soi_plan_t *)d->plan;
  return pl ? (void *)pl->stream : NULL;
}

void soi_enable_timing(soi_desc_t *d, int on)
{
  clear_err();
  soi_plan_t *pl = d ? (soi_plan_t *)d->plan : NULL;
  if (!pl) {
    set_err(SOI_ERR_STATE, "descriptor is not initialised (call init_soi_descriptor)");
    return;
  }
  if (on && !pl->ev[0][0]) {
    HIPCHK(hipSetDevice(pl->device));
    for (int c = 0; c < NTSLOT; ++c)
      for (int e = 0; e <= NSTAGE; ++e) HIPCHK(hipEventCreate(&pl->ev[c][e]));
  }
  pl->timing = on;
  pl->ncalls = 0;
fail:
  return;
}

int soi_stage_times(soi_desc_t *d, double *ms, int max_stages)
{
  clear_err();
  soi_plan_t *pl = d ? (soi_plan_t *)d->plan : NULL;
  if (!pl || !ms || pl->ncalls == 0) return 0;
  int n = max_stages < NSTAGE ? max_stages : NSTAGE;
  if (pl->filter_only_last && n > 1) n = 1;
  long first = pl->ncalls > NTSLOT ? pl->ncalls - NTSLOT : 0;
  int last_ev = pl->filter_only_last ? 1 : NSTAGE;
  HIPCHK(hipSetDevice(pl->device));
  HIPCHK(hipEventSynchronize(pl->ev[(pl->ncalls - 1) % NTSLOT][last_ev]));
  for (int e = 0; e < n; ++e) ms[e] = 0;
  for (long c = first; c < pl->ncalls; ++c) {
    hipEvent_t *ev = pl->ev[c % NTSLOT];
    for (int e = 0; e < n; ++e) {
      float t = 0;
      HIPCHK(hipEventElapsedTime(&t, ev[e], ev[e + 1]));
      ms[e] += t;
    }
  }
  for (int e = 0; e < n; ++e) ms[e] /= (double)(pl->ncalls - first);
  return n;
fail:
  return 0;
}

int soi_fft_variants(soi_desc_t *d)
{
  clear_err();
  soi_plan_t *pl = d ? (soi_plan_t *)d->plan : NULL;
  if (!pl) {
    set_err(SOI_ERR_STATE, "descriptor is not initialised (call init_soi_descriptor)");
    return 0;
  }
  return pl->nvar;
}

int soi_set_fft_variant(soi_desc_t *d, int v)
{
  clear_err();
  soi_plan_t *pl = d ? (soi_plan_t *)d->plan : NULL;
  if (!pl) {
    set_err(SOI_ERR_STATE, "descriptor is not initialised (call init_soi_descriptor)");
    return -1;
  }
  if (v < 0 || v >= pl->nvar) {
    set_err(SOI_ERR_ARGS, "FFT variant %d out of range [0, %d)", v, pl->nvar);
    return -1;
  }
  pl->var = v;
  if (pl->fft_reg) {
    pl->Tc = pl->vtab[v][0];
    pl->clay = pl->vtab[v][1];
    pl->Tr = pl->vtab[v][2];
  }
  return 0;
}

int soi_autotune(soi_desc_t *d, cfft_complex_t *scratch, int reps)
{
  clear_err();
  soi_plan_t *pl = d ? (soi_plan_t *)d->plan : NULL;
  if (!pl) {
    set_err(SOI_ERR_STATE, "descriptor is not initialised (call init_soi_descriptor)");
    return -1;
  }
  if (!scratch || !is_device_ptr(scratch)) {
    set_err(SOI_ERR_ARGS, "autotune needs a device scratch buffer of N elements");
    return -1;
  }
  if (reps < 1) reps = 3;
  const int was_timing = pl->timing;
  int best = 0;
  double best_ms = 1e300;
  for (int v = 0; v < pl->nvar; ++v) {
    if (soi_set_fft_variant(d, v)) return -1;
    compute_impl(d, scratch, NULL, 1, 0); /* warm-up (first launch sets kernel attributes) */
    if (soi_last_status()) return -1;
    soi_enable_timing(d, 1);
    for (int r = 0; r < reps; ++r) {
      compute_impl(d, scratch, NULL, 0, 0);
      if (soi_last_status()) return -1;
    }
    double ms[NSTAGE];
    int n = soi_stage_times(d, ms, NSTAGE);
    double t = n == NSTAGE ? ms[2] + ms[3] : 1e300;
    if (t < best_ms) best_ms = t, best = v;
  }
  soi_enable_timing(d, was_timing);
  if (soi_set_fft_variant(d, best)) return -1;
  return best;
}

int soi_filter_form(soi_desc_t *d)
{
  clear_err();
  soi_plan_t *pl = d ? (soi_plan_t *)d->plan : NULL;
  if (!pl) {
    set_err(SOI_ERR_STATE, "descriptor is not initialised");
    return -1;
  }
  return pl->fa.wt ? pl->fa.mfma : SOI_FILTER_VALU;
}

int soi_vlc_exponents(soi_desc_t *d, int *seg_max_exp, int *total_max_exp, int *send_counts)
{
  soi_plan_t *pl = (soi_plan_t *)d->plan;
  if (!pl || !pl->vlc_valid) return 0;
  (void)hipSetDevice(pl->device);
  (void)hipStreamSynchronize(pl->last_stream); /* the plan's copy of the last run has landed */
  if (seg_max_exp) memcpy(seg_max_exp, pl->h_vlc + 1, sizeof(int) * pl->S);
  if (total_max_exp) *total_max_exp = pl->h_vlc[0];
  if (send_counts) memcpy(send_counts, pl->h_vlc + 1 + pl->S, sizeof(int) * pl->S);
  return (int)pl->S;
}

/* Device-resident helpers for tests and the benchmark: seeded synthetic inputs. */
int soi_fill_uniform(void *x, long N, uint64_t seed, void *stream)
{
  int e = soik_gen_uniform((soi_c64_t *)x, N, seed, (hipStream_t)stream);
  if (!e) e = (int)hipStreamSynchronize((hipStream_t)stream);
  return e;
}

int soi_fill_tones(void *x, long N, const double *freqs, const double *phases, int ntones,
                   double amp, double noise_sigma, uint64_t seed, void *stream)
{
  double *df = NULL, *dp = NULL;
  int e = 0;
  if (ntones > 0) {
    if ((e = (int)hipMalloc((void **)&df, ntones * sizeof(double)))) return e;
    if ((e = (int)hipMalloc((void **)&dp, ntones * sizeof(double)))) {
      (void)hipFree(df);
      return e;
    }
    (void)hipMemcpy(df, freqs, ntones * sizeof(double), hipMemcpyHostToDevice);
    (void)hipMemcpy(dp, phases, ntones * sizeof(double), hipMemcpyHostToDevice);
  }
  e = soik_gen_tones((soi_c64_t *)x, N, df, dp, ntones, amp, noise_sigma, seed,
                     (hipStream_t)stream);
  if (!e) e = (int)hipStreamSynchronize((hipStream_t)stream);
  (void)hipFree(df);
  (void)hipFree(dp);
  return e;
}

/* ---- batching form: many initialised descriptors replayed as one HIP graph ---------------- */

#define SOI_BATCH_MAX_BRANCHES 16 /* independent capture branches at most */

struct soi_batch_s {
  int count;
  int device;
  hipStream_t stream;
  int nbranch;
  hipStream_t branch[SOI_BATCH_MAX_BRANCHES];
  hipEvent_t fork, join[SOI_BATCH_MAX_BRANCHES];
  hipGraph_t graph;
  hipGraphExec_t exec;
  int ngroups;
  void **tabs; /* device pointer tables of the groups (host array of ngroups entries) */
};

static int same_shape(const soi_desc_t *a, const soi_desc_t *b)
{
  return a->N == b->N && a->P == b->P && a->k == b->k && a->precision == b->precision &&
         a->n_mu == b->n_mu && a->d_mu == b->d_mu && a->B == b->B && a->tau == b->tau &&
         a->sigma == b->sigma && a->window_mode == b->window_mode;
}

/* Transforms that one batched launch can take together: same shape and the same kernel choices
 * (so that the first one's tables and launch shapes serve all of them). */
static int same_plan(const soi_desc_t *a, const soi_desc_t *b)
{
  const soi_plan_t *pa = (const soi_plan_t *)a->plan, *pb = (const soi_plan_t *)b->plan;
  return same_shape(a, b) && pa->var == pb->var && pa->tiled == pb->tiled &&
         pa->Tc == pb->Tc && pa->Tr == pb->Tr && pa->clay == pb->clay &&
         (pa->fa.wt != NULL) == (pb->fa.wt != NULL);
}

void soi_batch_destroy(soi_batch_t *b)
{
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (int r = 0; r < SOI_BATCH_MAX_BRANCHES; ++r) {
    if (b->branch[r]) (void)hipStreamDestroy(b->branch[r]);
    if (b->join[r]) (void)hipEventDestroy(b->join[r]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (b->tabs) {
    for (int g = 0; g < b->ngroups; ++g) (void)hipFree(b->tabs[g]);
    free(b->tabs);
  }
  free(b);
}

soi_batch_t *soi_batch_create(soi_desc_t **descs, cfft_complex_t **bufs, int count, int autotune)
{
  clear_err();
  soi_batch_t *b = NULL;
  int capturing = 0;
  unsigned char *tsave = NULL; /* the descriptors' timing flags, restored on every exit */
  if (!descs || !bufs || count < 1) {
    set_err(SOI_ERR_ARGS, "soi_batch_create needs count >= 1 descriptors and buffers");
    return NULL;
  }
  int device = -1;
  for (int i = 0; i < count; ++i) {
    soi_plan_t *pl = descs[i] ? (soi_plan_t *)descs[i]->plan : NULL;
    if (!pl) {
      set_err(SOI_ERR_STATE, "batch entry %d: descriptor is not initialised", i);
      return NULL;
    }
    if (descs[i]->use_vlc) {
      set_err(SOI_ERR_ARGS, "batch entry %d: use_vlc needs host round trips and cannot be "
              "captured in a graph", i);
      return NULL;
    }
    if (!bufs[i] || !is_device_ptr(bufs[i])) {
      set_err(SOI_ERR_ARGS, "batch entry %d: buffer must be a device pointer", i);
      return NULL;
    }
    if (device >= 0 && pl->device != device) {
      set_err(SOI_ERR_ARGS, "batch entry %d: all descriptors must belong to one device", i);
      return NULL;
    }
    device = pl->device;
  }
  HIPCHK(hipSetDevice(device));
  /* autotune: one timing sweep per distinct shape, on a scratch buffer */
  if (autotune) {
    for (int i = 0; i < count; ++i) {
      int j = 0;
      while (j < i && !same_shape(descs[j], descs[i])) ++j;
      if (j < i) {
        soi_plan_t *pj = (soi_plan_t *)descs[j]->plan;
        if (soi_set_fft_variant(descs[i], pj->var)) return NULL;
        continue;
      }
      soi_plan_t *pl = (soi_plan_t *)descs[i]->plan;
      size_t bytes = (size_t)pl->N * (pl->f32 ? sizeof(soi_c32_t) : sizeof(soi_c64_t));
      void *scratch = NULL;
      if (hipMalloc(&scratch, bytes) != hipSuccess) {
        (void)hipGetLastError();
        set_err(SOI_ERR_ALLOC, "autotune scratch of %zu bytes", bytes);
        return NULL;
      }
      (void)hipMemset(scratch, 0, bytes);
      int r = soi_autotune(descs[i], (cfft_complex_t *)scratch, 3);
      (void)hipFree(scratch);
      if (r < 0) return NULL;
    }
  }
  b = (soi_batch_t *)calloc(1, sizeof(*b));
  if (!b) {
    set_err(SOI_ERR_ALLOC, "host allocation failed");
    return NULL;
  }
  b->count = count;
  b->device = device;
  HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  /* SOI_BATCH_BRANCHES: concurrent capture branches (default 4) */
  const char *eb = getenv("SOI_BATCH_BRANCHES");
  b->nbranch = eb ? atoi(eb) : 4;
  if (b->nbranch < 1) b->nbranch = 1;
  if (b->nbranch > SOI_BATCH_MAX_BRANCHES) b->nbranch = SOI_BATCH_MAX_BRANCHES;
  for (int r = 0; r < b->nbranch; ++r) {
    HIPCHK(hipStreamCreateWithFlags(&b->branch[r], hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&b->join[r], hipEventDisableTiming));
  }
  /* timing events would become graph nodes: capture with timing off */
  tsave = (unsigned char *)malloc((size_t)count);
  if (!tsave) {
    set_err(SOI_ERR_ALLOC, "host allocation failed");
    goto fail_quiet;
  }
  for (int i = 0; i < count; ++i) {
    tsave[i] = (unsigned char)((soi_plan_t *)descs[i]->plan)->timing;
    ((soi_plan_t *)descs[i]->plan)->timing = 0;
  }
  /* groups of transforms that run as one launch per kernel (blockIdx.y = member): member
   * pointer tables [x][alpha_tilde][tile buffer] in device memory */
  int *grp = (int *)malloc(sizeof(int) * count), *gsize = (int *)calloc(count, sizeof(int));
  int *gfirst = (int *)malloc(sizeof(int) * count);
  void **htab = (void **)malloc(sizeof(void *) * 3 * count);
  double *gwork = (double *)calloc(count, sizeof(double)), *load = NULL;
  if (!grp || !gsize || !gfirst || !htab || !gwork) {
    free(grp), free(gsize), free(gfirst), free(htab), free(gwork);
    set_err(SOI_ERR_ALLOC, "host allocation failed");
    goto fail_quiet;
  }
  b->ngroups = 0;
  for (int i = 0; i < count; ++i) {
    int g = 0;
    while (g < b->ngroups && !same_plan(descs[gfirst[g]], descs[i])) ++g;
    if (g == b->ngroups) gfirst[b->ngroups++] = i;
    grp[i] = g;
    gsize[g]++;
    gwork[g] += (double)descs[i]->N * log2((double)descs[i]->N);
  }
  b->tabs = (void **)calloc(b->ngroups, sizeof(void *));
  load = (double *)calloc(b->nbranch, sizeof(double));
  if (!b->tabs || !load) {
    free(grp), free(gsize), free(gfirst), free(htab), free(gwork), free(load);
    set_err(SOI_ERR_ALLOC, "host allocation failed");
    goto fail_quiet;
  }
  for (int g = 0; g < b->ngroups; ++g) {
    const int n = gsize[g];
    for (int i = 0, m = 0; i < count; ++i) {
      if (grp[i] != g) continue;
      htab[m] = bufs[i];
      htab[n + m] = ((soi_plan_t *)descs[i]->plan)->at;
      htab[2 * n + m] = ((soi_plan_t *)descs[i]->plan)->tbuf;
      ++m;
    }
    hipError_t e = hipMalloc(&b->tabs[g], sizeof(void *) * 3 * n);
    if (e == hipSuccess)
      e = hipMemcpy(b->tabs[g], htab, sizeof(void *) * 3 * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      free(grp), free(gsize), free(gfirst), free(htab), free(gwork), free(load);
      set_err(SOI_ERR_HIP, "batch tables: %s", hipGetErrorString(e));
      goto fail_quiet;
    }
  }
  free(htab);
  {
    hipError_t ce = hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed);
    if (ce == hipSuccess) {
      capturing = 1;
      ce = hipEventRecord(b->fork, b->stream);
    }
    for (int r = 0; r < b->nbranch && ce == hipSuccess; ++r)
      ce = hipStreamWaitEvent(b->branch[r], b->fork, 0);
    if (ce != hipSuccess) {
      free(grp), free(gsize), free(gfirst), free(gwork), free(load);
      set_err(SOI_ERR_HIP, "batch capture: %s", hipGetErrorString(ce));
      goto fail_quiet;
    }
  }
  /* groups in decreasing work, each on the least loaded branch */
  for (int k = 0; k < b->ngroups; ++k) {
    int g = -1;
    for (int h = 0; h < b->ngroups; ++h)
      if (gwork[h] >= 0 && (g < 0 || gwork[h] > gwork[g])) g = h;
    int r = 0;
    for (int q = 1; q < b->nbranch; ++q)
      if (load[q] < load[r]) r = q;
    load[r] += gwork[g];
    gwork[g] = -1;
    const int n = gsize[g];
    void *const *tab = (void *const *)b->tabs[g];
    SoiBatch bt = {tab, tab + n, tab + 2 * n, n};
    soi_desc_t *d0 = descs[gfirst[g]];
    soi_plan_t *p0 = (soi_plan_t *)d0->plan;
    if (run_filter(d0, p0, bufs[gfirst[g]], bt, b->branch[r]) ||
        run_fft(d0, p0, bufs[gfirst[g]], bt, b->branch[r])) {
      free(grp), free(gsize), free(gfirst), free(gwork), free(load);
      goto fail_quiet;
    }
  }
  free(grp), free(gsize), free(gfirst), free(gwork), free(load);
  for (int r = 0; r < b->nbranch; ++r) {
    HIPCHK(hipEventRecord(b->join[r], b->branch[r]));
    HIPCHK(hipStreamWaitEvent(b->stream, b->join[r], 0));
  }
  capturing = 0;
  HIPCHK(hipStreamEndCapture(b->stream, &b->graph));
  HIPCHK(hipGraphInstantiate(&b->exec, b->graph, NULL, NULL, 0));
  for (int i = 0; i < count; ++i) ((soi_plan_t *)descs[i]->plan)->timing = tsave[i];
  free(tsave);
  return b;
fail:
fail_quiet:
  if (capturing) {
    hipGraph_t g = NULL;
    (void)hipStreamEndCapture(b->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
  }
  if (tsave)
    for (int i = 0; i < count; ++i) ((soi_plan_t *)descs[i]->plan)->timing = tsave[i];
  free(tsave);
  soi_batch_destroy(b);
  return NULL;
}

int soi_batch_run(soi_batch_t *b, void *stream)
{
  clear_err();
  if (!b || !b->exec) {
    set_err(SOI_ERR_STATE, "batch is not created");
    return SOI_ERR_STATE;
  }
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipGraphLaunch(b->exec, stream ? (hipStream_t)stream : b->stream));
  return 0;
fail:
  return soi_last_status();
}

void *soi_batch_stream(soi_batch_t *b) { return b ? (void *)b->stream : NULL; }
