/*
 * soi.h -- public C ABI of the MI355X SOI FFT (drop-in for IntelLabs/SOI_FFT's soi.h).
 *
 * The descriptor create/compute/destroy API keeps the reference's names, argument meaning and
 * caller-set fields (reference soi.h:236-280).  One process drives one GPU; the P MPI ranks of
 * the reference become P virtual ranks on the device ("P ranks -> P virtual segments"): the
 * caller sets N, P (and k), leaves rank = 0 and passes the whole N-element input.  On return
 * the buffer holds all S = k*P segments of the N-point forward DFT in natural order, exactly
 * the concatenation of what the reference's ranks 0..P-1 hold after compute_soi.
 *
 * Buffers may be device pointers (HBM-resident, the fast path) or host pointers (copied in and
 * out around the GPU work); hipPointerGetAttributes tells them apart.
 *
 * Errors: the reference prints and exit()s (parallel_filter_subsampling.cpp:89-94, 537-548,
 * pfft.c:157-183).  Here every entry point records a message, and soi_last_error() returns it;
 * soi_last_status() returns its code (0 = success).  Nothing exits the process.
 */
#ifndef SOI_AMD_SOI_H_
#define SOI_AMD_SOI_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* MPI is not needed on one GPU.  When the caller's translation unit has included mpi.h the
 * real MPI_Comm is used; otherwise a placeholder integer type is declared.  The descriptor
 * keeps the handle in an 8-byte slot either way, so the layout does not depend on it. */
#ifndef MPI_COMM_WORLD
typedef int MPI_Comm;
#define SOI_NO_MPI_ 1
#define MPI_COMM_WORLD ((MPI_Comm)0x44000000)
#endif

typedef size_t cfft_size_t;                 /* reference soi.h:220 */

typedef struct _complex_struct_t {          /* reference soi.h:222-225 */
  double re;
  double im;
} complex_struct_t;

/* reference intrinsic.h:4-5: PRECISION 2 (the default) is double precision, PRECISION 1 single.
 * reference soi.h:227: cfft_complex_t is C99 double complex (C++ callers see a layout-identical
 * struct).  A caller built with -DPRECISION=1 gets float complex buffers, and
 * set_default_soi_descriptor selects the FP32 form for it (below). */
#ifndef PRECISION
#define PRECISION 2
#endif
#if PRECISION == 1
#ifdef __cplusplus
typedef struct _complex_struct32_t {
  float re;
  float im;
} cfft_complex_t;
#else
#include <complex.h>
typedef float complex cfft_complex_t;
#endif
#else
#ifdef __cplusplus
typedef complex_struct_t cfft_complex_t;
#else
#include <complex.h>
typedef double complex cfft_complex_t;
#endif
#endif

/* Precision / form selector (not in the reference, which is FP64 only).
 *   SOI_PREC_FP64      plain FP64 path, the drop-in (default)
 *   SOI_PREC_FP32      precision form: complex64 input, output and alpha_tilde (a caller that
 *                      preallocates alpha_tilde sizes it S*M_hat complex64 values); FP32 filter
 *                      contraction and FFT arithmetic with windows and twiddles generated in
 *                      FP64 and rounded once; W_inv applied in FP64.  Not combinable with
 *                      use_vlc (the VLC words encode FP64 values). */
enum { SOI_PREC_FP64 = 0, SOI_PREC_FP32 = 1 };

/* Window evaluation (not in the reference).
 *   SOI_WINDOW_AUTO       exact windows with real filter taps where they agree with the
 *                         reference's window values to far below 1e-12 in the output
 *                         ((n_mu,d_mu) = (5,4) and B >= 60), the reference's values otherwise
 *   SOI_WINDOW_REFERENCE  the reference's window values: t formed in double as pfft.c:52 does
 *   SOI_WINDOW_EXACT      t from its exact rational value (more accurate than the reference) */
enum { SOI_WINDOW_AUTO = 0, SOI_WINDOW_REFERENCE = 1, SOI_WINDOW_EXACT = 2 };

/* Status codes returned by soi_last_status(). */
enum {
  SOI_OK = 0,
  SOI_ERR_ARGS = 1,          /* invalid or unsupported descriptor parameters */
  SOI_ERR_TOO_SMALL = 2,     /* M/P < B (parallel_filter_subsampling.cpp:89-94) */
  SOI_ERR_UNSUPPORTED_MU = 3,/* (n_mu,d_mu) not in {(5,4),(8,7)} (:537-548) */
  SOI_ERR_ALLOC = 4,         /* device allocation failed (pfft.c:157-183) */
  SOI_ERR_HIP = 5,           /* a HIP runtime call failed */
  SOI_ERR_STATE = 6          /* descriptor used before init or after free */
};

/* Descriptor, reference soi.h:236-270.  Field names and meanings are kept; DFTI handles,
 * w_dup and MPI requests are replaced by one opaque GPU plan. */
typedef struct {
  union { MPI_Comm comm; int64_t comm_slot_; };  /* unused on one GPU */
  int P;                      /* number of (virtual) ranks */
  int rank;                   /* must be 0 on one GPU */
  cfft_size_t k;              /* segments per rank */
  cfft_size_t N;              /* global vector length */
  cfft_complex_t *W_inv;      /* device: inverse frequency window, M values */
  cfft_complex_t *w;          /* device: time window, n_mu*B*S values, layout [theta][b][i] */
  void *w_dup;                /* device: real filter taps used by the kernels */
  cfft_complex_t *gamma_tilde;/* device: FFT work buffer, S*M_hat (caller may preallocate) */
  cfft_complex_t *alpha_tilde;/* device: filtered and subsampled data [S][M_hat] (ditto) */
  cfft_complex_t *beta_tilde; /* kept for layout compatibility; never allocated */
  cfft_complex_t *alpha_ghost;/* unused: the halo is circular indexing in HBM */
  int *delta, *epsilon;       /* device: VLC word streams (allocated when use_vlc) */

  int n_mu;
  int d_mu;
  cfft_size_t B;
  double tau;
  double sigma;

  void *plan;                 /* opaque GPU plan (replaces desc_dft_s / desc_dft_m_hat) */
  void *sendRequests_, *recvRequests_;  /* layout slots for the reference's request arrays */
  int use_vlc;                /* variable length compression of the exchanged data */
  int *segmentBoundaries;     /* host, P+1 entries: first segment of each rank (with use_vlc
                                 the cost-balanced boundaries of the last run, pfft.c:406-429,
                                 valid once that run's stream has completed) */
  double comm_to_comp_cost_ratio;
  int precision;              /* SOI_PREC_FP64 (default) or SOI_PREC_FP32 */
  int window_mode;            /* SOI_WINDOW_AUTO (default), _REFERENCE or _EXACT */
} soi_desc_t;

/* ---- reference API (soi.h:272-280) ---------------------------------------------------- */

/* pfft.c:84-138: n_mu=5, d_mu=4, tau=928/1024, sigma=373.7314, B=72, use_vlc=0,
 * comm_to_comp_cost_ratio=1; nulls the reusable buffers.  Also P=1, rank=0, k=1, N=0. */
void set_default_soi_descriptor(soi_desc_t *desc);

/* set_default_soi_descriptor with the precision field set as well (no counterpart in the
 * reference).  A caller built with -DPRECISION=1 reaches it through set_default_soi_descriptor:
 * the reference's single-precision build (intrinsic.h:34-56) maps onto the FP32 form. */
void soi_set_default_descriptor_precision(soi_desc_t *desc, int precision);
#if PRECISION == 1
#define set_default_soi_descriptor(desc) soi_set_default_descriptor_precision((desc), SOI_PREC_FP32)
#endif

/* pfft.c:140-232: sets comm and k, segment boundaries p*k, allocates device buffers (unless
 * preallocated), builds windows and taps on the GPU in FP64, plans the FFTs.  The caller must
 * preset N and P (and rank = 0). */
void init_soi_descriptor(soi_desc_t *desc, MPI_Comm comm, cfft_size_t k);

/* pfft.c:276-721: in-place forward DFT of alpha_dt (N elements, device or host pointer).
 * With precision == SOI_PREC_FP32 the buffer holds N complex64 values instead (passed through
 * the same cfft_complex_t * parameter). */
void compute_soi(soi_desc_t *d, cfft_complex_t *alpha_dt);

/* pfft.c:234-267 */
void free_soi_descriptor(soi_desc_t *d);

/* parallel_filter_subsampling.cpp:534-550: Step 1 only.  Reads alpha_dt (N elements) and
 * leaves d->alpha_tilde = [S][M_hat] (device). */
void parallel_filter_subsampling(soi_desc_t *desc, cfft_complex_t *alpha_dt);

/* ---- additions for a checked, stream-aware C ABI ----------------------------------------- */

/* Message and code of the last error on this thread ("" and 0 after a successful call). */
const char *soi_last_error(void);
int soi_last_status(void);

/* compute_soi on a caller-chosen HIP stream (hipStream_t passed as void*; NULL = the
 * descriptor's own stream) without a final synchronisation: returns after enqueueing when
 * alpha_dt is a device pointer, with use_vlc too (the VLC plan is computed on the device and
 * copied to segmentBoundaries asynchronously). */
void compute_soi_async(soi_desc_t *d, cfft_complex_t *alpha_dt, void *stream);

/* Stream the descriptor's kernels run on (hipStream_t as void*). */
void *soi_stream(soi_desc_t *d);

/* Per-stage GPU times in milliseconds (filter + S-point DFT, vlc, fft pass 1, fft pass 2 +
 * demodulation), from HIP events recorded on the stream each call runs on.  Enabling timing
 * resets the record; soi_stage_times waits for the last recorded call and returns the mean over
 * the last min(calls, 64) calls since then (the filter stage only after
 * parallel_filter_subsampling).  Returns the number of stages written, 0 if none was recorded. */
void soi_enable_timing(soi_desc_t *d, int on);
int soi_stage_times(soi_desc_t *d, double *ms, int max_stages);

/* Form of the plan's filter contraction (not in the reference): SOI_FILTER_VALU (FP64 or FP32
 * FMAs on the vector units), SOI_FILTER_MFMA_BF16X3 (complex64, split-bf16 MFMA),
 * SOI_FILTER_MFMA_F32 (complex64, v_mfma_f32_16x16x4_f32) or SOI_FILTER_MFMA_F64 (complex128,
 * v_mfma_f64_16x16x4_f64); -1 for a descriptor that is not initialised. */
enum { SOI_FILTER_VALU = 0, SOI_FILTER_MFMA_BF16X3 = 1, SOI_FILTER_MFMA_F32 = 2,
       SOI_FILTER_MFMA_F64 = 3 };
int soi_filter_form(soi_desc_t *d);

/* Per-segment maximum exponents and their global maximum from the last VLC run
 * (pfft.c:319-343), and the per-segment send counts in doubles (pfft.c:404). */
int soi_vlc_exponents(soi_desc_t *d, int *seg_max_exp, int *total_max_exp, int *send_counts);

/* ---- FFT tuning ----------------------------------------------------------------------------
 * The segment FFT of a plan has one or more equivalent kernel configurations (tile widths and
 * register layouts).  soi_fft_variants returns their number, soi_set_fft_variant selects one,
 * and soi_autotune times each on `scratch` (a device buffer of N elements whose contents are
 * overwritten) over `reps` calls, keeps the fastest and returns its index (-1 on error). */
int soi_fft_variants(soi_desc_t *d);
int soi_set_fft_variant(soi_desc_t *d, int variant);
int soi_autotune(soi_desc_t *d, cfft_complex_t *scratch, int reps);

/* ---- batching form ----------------------------------------------------------------------
 * Many independent transforms (initialised descriptors of any shapes and either precision,
 * without use_vlc), each with its own device buffer, captured once into one HIP graph:
 * transforms of the same shape run as one launch per kernel, and those groups run on concurrent
 * graph branches; soi_batch_run replays it with one launch (in place on the same buffers, which
 * the caller refills between runs).  With autotune != 0 each distinct shape is tuned once
 * (soi_autotune) before capture.  SOI_BATCH_BRANCHES (1..16, default 4) sets the branch count.
 * Returns NULL on error (soi_last_error). */
typedef struct soi_batch_s soi_batch_t;
soi_batch_t *soi_batch_create(soi_desc_t **descs, cfft_complex_t **bufs, int count, int autotune);
int soi_batch_run(soi_batch_t *b, void *stream);   /* NULL = the batch's own stream */
void *soi_batch_stream(soi_batch_t *b);
void soi_batch_destroy(soi_batch_t *b);

/* ---- test-harness entry points (reference soi.h:282-288, 306, 314-315) --------------------
 * Host buffers; input kinds 0 (rectangle), 1 (e^{i idx}) and 3 (Dirichlet kernel) with their
 * closed-form transforms (reference input.c:218-290).  Kinds 2, 4 and 5 need MKL VML/CDFT or
 * data files: they set SOI_ERR_ARGS, leave the buffer untouched and compute_* return NaN.
 * reference_output of kind 3 is the reference's value, 1/N of the exact transform.  On one GPU
 * offset = 0 and localLen = globalLen cover the whole transform. */
void populate_input(cfft_complex_t *input, size_t localLen, size_t offset, size_t globalLen,
                    int kind);
cfft_complex_t reference_output(size_t idx, size_t globalLen, int kind, size_t offset);
double compute_snr(cfft_complex_t *output, size_t localLen, size_t offset, size_t globalLen,
                   int kind, soi_desc_t *d);
double compute_normalized_inf_norm(cfft_complex_t *output, size_t localLen, size_t offset,
                                   size_t globalLen, int kind);
double get_cpu_freq(void); /* TSC ticks per second */
#define SOI_MAX_THREADS 512 /* reference soi.h MAX_THREADS */
/* per-thread cycle counters of the reference's CPU loops; the GPU path leaves them zero */
extern unsigned long long load_imbalance_times[SOI_MAX_THREADS];

#ifdef __cplusplus
}
#endif

#endif /* SOI_AMD_SOI_H_ */
