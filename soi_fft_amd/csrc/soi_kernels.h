/*
 * soi_kernels.h -- internal C ABI between the C host layer (pfft.c, compress_host.c) and the
 * HIP kernels (soi_kernels.hip).  Device pointers only; every launcher enqueues on `st` and
 * returns a hipError_t value (0 = success).
 */
#ifndef SOI_AMD_KERNELS_H_
#define SOI_AMD_KERNELS_H_

#include <stdint.h>

#ifdef __cplusplus
#include <hip/hip_runtime.h>
typedef double2 soi_c64_t;
typedef float2 soi_c32_t;
extern "C" {
#else
#include <hip/hip_runtime_api.h>
typedef struct { double x, y; } soi_c64_t;
typedef struct { float x, y; } soi_c32_t;
#endif

/* Batched launch: n transforms of one shape in one launch, blockIdx.y = transform.  The tables
 * are device arrays of the per-transform buffers (in/out vector, alpha_tilde, pass-1 tile
 * buffer).  n = 0: one transform on the pointers passed directly. */
typedef struct SoiBatch {
  void *const *x;
  void *const *at;
  void *const *tb;
  int n;
} SoiBatch;

typedef struct SoiFilterArgs {
  const void *x;      /* input [M][S], complex128 (f32 = 0) or complex64 (f32 = 1) */
  void *at;           /* output alpha_tilde [S][M_hat], same element type */
  const void *taps;   /* [n_mu][S][NP*BP], real (ct = 0) or complex (ct = 1), double or float */
  const void *post;   /* [n_mu][S] post phase, complex of the same precision */
  const void *twS;    /* S-point twiddles, complex of the same precision */
  const void *wt;     /* tap operands of an MFMA form (NULL: VALU form) */
  long M, Mh, nJ;
  int S, n_mu, d_mu, BP, NP, NT, U, W, ct, f32;
  int mfma;           /* MFMA form of wt: 1 = split-bf16 pieces, 2 = FP32, 3 = FP64 (complex128) */
  SoiBatch b;         /* batched launch tables (x -> b.x[t], at -> b.at[t]) */
} SoiFilterArgs;

int soik_make_tables(void *taps, soi_c64_t *post, soi_c64_t *w_full, soi_c64_t *winv, int S,
                     int n_mu, int d_mu, int B, int B_pad, double tau, double sigma, long N,
                     long M, long Mh, int ref, int ct, hipStream_t st);
int soik_filter_shape(int S, int n_mu, int d_mu, int B, int ct, int *NP, int *BP, int *NT,
                      int *U, int *W);
size_t soik_filter_lds_bytes(const SoiFilterArgs *a);
int soik_make_twiddles(soi_c64_t *tw, long n, long L, long mult, hipStream_t st);
int soik_filter_sdft(const SoiFilterArgs *a, hipStream_t st);
int soik_filter_mfma_supported(int S, int n_mu, int d_mu, int ct, int BPAD, int mode);
size_t soik_filter_mfma_table_bytes(int S, int BPAD, int mode);
int soik_filter_mfma_table(const double *taps, void *wt, int S, int BPAD, int mode,
                           hipStream_t st);
int soik_fft_single(const soi_c64_t *at, soi_c64_t *X, const soi_c64_t *tw,
                    const soi_c64_t *winv, long L, long M, int S, int count, SoiBatch b, hipStream_t st);
int soik_fft_col(soi_c64_t *data, const soi_c64_t *tw1, const soi_c64_t *thi,
                 const soi_c64_t *tlo, int n1, int n2, int Tc, int B1, long Mh, int S,
                 SoiBatch b, hipStream_t st);
int soik_fft_row(const soi_c64_t *data, soi_c64_t *X, const soi_c64_t *tw2,
                 const soi_c64_t *winv, int n1, int n2, int Tr, long Mh, long M, int S,
                 SoiBatch b, hipStream_t st);
int soik_fft_single_f32(const soi_c32_t *at, soi_c32_t *X, const soi_c32_t *tw,
                        const soi_c64_t *winv, long L, long M, int S, int count, SoiBatch b, hipStream_t st);
int soik_fft_col_f32(soi_c32_t *data, const soi_c32_t *tw1, const soi_c32_t *thi,
                     const soi_c32_t *tlo, int n1, int n2, int Tc, int B1, long Mh, int S,
                     SoiBatch b, hipStream_t st);
int soik_fft_row_f32(const soi_c32_t *data, soi_c32_t *X, const soi_c32_t *tw2,
                     const soi_c64_t *winv, int n1, int n2, int Tr, long Mh, long M, int S,
                     SoiBatch b, hipStream_t st);
int soik_fft_reg_supported(int L, int rows);
int soik_fft_cols_short_tile(int L);
int soik_fft_cols_reg(soi_c64_t *data, const soi_c64_t *tw1, const soi_c64_t *thi,
                      const soi_c64_t *tlo, int n1, int n2, int lb1, long Mh, int S, int T,
                      int layout, soi_c64_t *tout, SoiBatch b, hipStream_t st);
int soik_fft_rows_reg(const soi_c64_t *data, soi_c64_t *X, const soi_c64_t *tw2,
                      const soi_c64_t *winv, int n1, int n2, long Mh, long M, int S, int T,
                      SoiBatch b, hipStream_t st);
int soik_fft_cols_reg_f32(soi_c32_t *data, const soi_c32_t *tw1, const soi_c32_t *thi,
                          const soi_c32_t *tlo, int n1, int n2, int lb1, long Mh, int S, int T,
                          int layout, soi_c32_t *tout, SoiBatch b, hipStream_t st);
/* pass 2 over pass 1's 8-row tiles (tout of the column pass) */
int soik_fft_rows_tiled_supported(int n2, int f32);
int soik_fft_rows_tiled(const soi_c64_t *data, soi_c64_t *X, const soi_c64_t *tw2,
                        const soi_c64_t *winv, int n1, int n2, long Mh, long M, int S,
                        SoiBatch b, hipStream_t st);
int soik_fft_rows_tiled_f32(const soi_c32_t *data, soi_c32_t *X, const soi_c32_t *tw2,
                            const soi_c64_t *winv, int n1, int n2, long Mh, long M, int S,
                            SoiBatch b, hipStream_t st);
int soik_fft_rows_reg_f32(const soi_c32_t *data, soi_c32_t *X, const soi_c32_t *tw2,
                          const soi_c64_t *winv, int n1, int n2, long Mh, long M, int S, int T,
                          SoiBatch b, hipStream_t st);
int soik_max_exponent(const double *x, int len, int nchunk, int *out, hipStream_t st);
int soik_compress(int *out, long stride, const double *x, int len, int nchunk, int e_max,
                  const int *emi, int *nwords, const int *e_max_dev, hipStream_t st);
int soik_decompress(double *out, const int *in, long stride, int len, int nchunk, int e_max,
                    const int *emi, const int *e_max_dev, hipStream_t st);
/* VLC planning on the device: out = [total, seg_max[S], send_counts[S], boundaries[P+1]] */
int soik_vlc_plan(const int *exps, int S, int P, int k, long l, long Mh, double ratio, int *emi,
                  int *out, double *pcs, long *upre, int *nwords, hipStream_t st);
int soik_compress_flat(int *out, long stride, const double *x, int len, int nchunk,
                       const int *e_max_dev, const int *emi, const long *upre, hipStream_t st);
int soik_decompress_flat(double *out, const int *in, long stride, int len, int nchunk,
                         const int *e_max_dev, const int *emi, hipStream_t st);
int soik_gen_uniform(soi_c64_t *x, long N, uint64_t seed, hipStream_t st);
int soik_gen_tones(soi_c64_t *x, long N, const double *tf, const double *tph, int ntones,
                   double amp, double noise_sigma, uint64_t seed, hipStream_t st);
int soik_f64_to_f32(const soi_c64_t *in, soi_c32_t *out, long n, hipStream_t st);

#ifdef __cplusplus
}
#endif

#endif
