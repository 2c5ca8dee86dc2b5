/*
 * compress.h -- VLC codec C ABI (drop-in for the reference's compress.h:3-10).
 *
 * Same names, arguments and results as the reference.  Pointers may be host or device
 * memory.  The codec is bit-exact with the reference compress.c: the word stream of
 * compress() (16 interleaved 32-bit lanes, round-half-even body, round-half-away tail) and the
 * doubles decompress() returns for the 16-lane body.  decompress() decodes the scalar tail as
 * the exact inverse of compress(); see the reference-quirk note in the implementation.
 */
#ifndef SOI_AMD_COMPRESS_H_
#define SOI_AMD_COMPRESS_H_

#ifdef __cplusplus
extern "C" {
#endif

/* compress.c:11-131: max over len doubles of the biased exponent field, minus 1023. */
int max_exponent(const double *x, int len);

/* compress.c:205-372: returns the number of 32-bit words written to out. */
int compress(int *out, const double *in, int len, int e_max, int e_max_i, int print);

/* compress.c:428-544: refIn is only read by the reference's debug builds; ignored. */
void decompress(double *out, const int *in, int len, int e_max, int e_max_i, const double *refIn);

/* Batched device forms used by compute_soi and by callers that keep data in HBM.
 *   x: nchunk chunks of len doubles, chunk c at x + c*len.
 *   exps: nchunk ints.  e_max_i: per-chunk segment exponent; out words for chunk c at
 *   out + c*stride_words; nwords[c] receives compress()'s return value for chunk c.
 *   All pointers are device pointers; stream is a hipStream_t (NULL = default stream). */
int soi_max_exponent_batch(const double *x, int len, int nchunk, int *exps, void *stream);
int soi_compress_batch(int *out, long stride_words, const double *x, int len, int nchunk,
                       int e_max, const int *e_max_i, int *nwords, void *stream);
int soi_decompress_batch(double *out, const int *in, long stride_words, int len, int nchunk,
                         int e_max, const int *e_max_i, void *stream);

/* No counterpart in the reference.  An allocation-free compress(), for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the batched form of compress() over nchunk chunks of len doubles (chunk c
 * at x + c*len, its words at out + c*stride_words), x and out device pointers, e_max_i and
 * nwords host arrays of nchunk ints (the segment exponents in, compress()'s return value per
 * chunk out).  The exponents and counts pass through a static device scratch of at most 4096
 * chunks (more is rejected with status 1); nothing is allocated.  Synchronises the stream
 * (hipStream_t, NULL = default) before returning.  Returns 0 or a soi.h status code
 * (soi_last_error describes it).  Calls that overlap on one device share the scratch and race. */
int soi_compress_noalloc(int *out, long stride_words, const double *x, int len, int nchunk,
                         int e_max, const int *e_max_i, int *nwords, void *stream);

#ifdef __cplusplus
}
#endif

#endif
