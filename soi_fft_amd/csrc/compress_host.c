/*
 * compress_host.c -- the reference's VLC codec API (compress.h:3-10) over the GPU codec kernels.
 *
 * max_exponent / compress / decompress keep the reference's signatures (compress.c:11, 205,
 * 428).  Host pointers are staged through device memory; device pointers are used in place.
 * The codec itself is in soi_kernels.hip (k_max_exponent, k_compress, k_decompress).
 */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../include/compress.h"
#include "../../include/soi.h"
#include "soi_kernels.h"

static int dev_ptr(const void *p)
{
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

int soi_max_exponent_batch(const double *x, int len, int nchunk, int *exps, void *stream)
{
  return soik_max_exponent(x, len, nchunk, exps, (hipStream_t)stream);
}

int soi_compress_batch(int *out, long stride_words, const double *x, int len, int nchunk,
                       int e_max, const int *e_max_i, int *nwords, void *stream)
{
  return soik_compress(out, stride_words, x, len, nchunk, e_max, e_max_i, nwords, NULL,
                       (hipStream_t)stream);
}

int soi_decompress_batch(double *out, const int *in, long stride_words, int len, int nchunk,
                         int e_max, const int *e_max_i, void *stream)
{
  return soik_decompress(out, in, stride_words, len, nchunk, e_max, e_max_i, NULL,
                         (hipStream_t)stream);
}

/* words compress() writes for len doubles: vector rows then the scalar tail (compress.c:205-372) */
static long max_words(int len) { return (long)len * 56 / 32 + 32; }

int max_exponent(const double *x, int len)
{
  int r = INT32_MIN;
  double *dx = NULL;
  int *dr = NULL;
  if (len <= 0) return (int)((unsigned)INT32_MIN - 1023u);
  int dev = dev_ptr(x);
  if (hipMalloc((void **)&dr, sizeof(int)) != hipSuccess) return r;
  const double *src = x;
  if (!dev) {
    if (hipMalloc((void **)&dx, (size_t)len * sizeof(double)) != hipSuccess) goto done;
    (void)hipMemcpy(dx, x, (size_t)len * sizeof(double), hipMemcpyHostToDevice);
    src = dx;
  }
  if (soik_max_exponent(src, len, 1, dr, 0) == 0)
    (void)hipMemcpy(&r, dr, sizeof(int), hipMemcpyDeviceToHost);
done:
  (void)hipFree(dx);
  (void)hipFree(dr);
  return r;
}

int compress(int *out, const double *in, int len, int e_max, int e_max_i, int print)
{
  (void)print;
  int nw = 0;
  int *demi = NULL, *dnw = NULL, *dout = NULL;
  double *din = NULL;
  int dev_in = dev_ptr(in), dev_out = dev_ptr(out);
  if (hipMalloc((void **)&demi, sizeof(int)) != hipSuccess) return 0;
  if (hipMalloc((void **)&dnw, sizeof(int)) != hipSuccess) goto done;
  (void)hipMemcpy(demi, &e_max_i, sizeof(int), hipMemcpyHostToDevice);
  const double *src = in;
  int *dst = out;
  if (!dev_in) {
    if (hipMalloc((void **)&din, (size_t)len * sizeof(double) + 8) != hipSuccess) goto done;
    (void)hipMemcpy(din, in, (size_t)len * sizeof(double), hipMemcpyHostToDevice);
    src = din;
  }
  if (!dev_out) {
    if (hipMalloc((void **)&dout, (size_t)max_words(len) * sizeof(int)) != hipSuccess) goto done;
    dst = dout;
  }
  if (soik_compress(dst, max_words(len), src, len, 1, e_max, demi, dnw, NULL, 0) == 0) {
    (void)hipMemcpy(&nw, dnw, sizeof(int), hipMemcpyDeviceToHost);
    if (!dev_out && nw > 0) (void)hipMemcpy(out, dout, (size_t)nw * sizeof(int), hipMemcpyDeviceToHost);
  }
done:
  (void)hipFree(demi);
  (void)hipFree(dnw);
  (void)hipFree(din);
  (void)hipFree(dout);
  return nw;
}

void decompress(double *out, const int *in, int len, int e_max, int e_max_i, const double *refIn)
{
  (void)refIn;
  int *demi = NULL, *din = NULL;
  double *dout = NULL;
  int dev_in = dev_ptr(in), dev_out = dev_ptr(out);
  if (hipMalloc((void **)&demi, sizeof(int)) != hipSuccess) return;
  (void)hipMemcpy(demi, &e_max_i, sizeof(int), hipMemcpyHostToDevice);
  const int *src = in;
  double *dst = out;
  if (!dev_in) {
    /* the caller's stream holds exactly compress()'s word count; copy what it can hold */
    long nw = max_words(len);
    if (hipMalloc((void **)&din, (size_t)nw * sizeof(int)) != hipSuccess) goto done;
    (void)hipMemset(din, 0, (size_t)nw * sizeof(int));
    /* number of words compress() produced for these parameters */
    int dl = e_max - e_max_i;
    int i1 = 24 - dl, i2 = (55 - dl) < 32 ? (55 - dl) : 32;
    long used = 0;
    if (i2 > 0) {
      long fb = (i1 > 0 ? i1 : 0) + i2;
      long nvf = len / 16, nt = len - nvf * 16;
      used = (nvf * fb + 31) / 32 * 16 + (nt * fb + 31) / 32;
    }
    if (used > 0) (void)hipMemcpy(din, in, (size_t)used * sizeof(int), hipMemcpyHostToDevice);
    src = din;
  }
  if (!dev_out) {
    if (hipMalloc((void **)&dout, (size_t)len * sizeof(double) + 8) != hipSuccess) goto done;
    dst = dout;
  }
  if (soik_decompress(dst, src, max_words(len), len, 1, e_max, demi, NULL, 0) == 0 && !dev_out)
    (void)hipMemcpy(out, dout, (size_t)len * sizeof(double), hipMemcpyDeviceToHost);
  else
    (void)hipDeviceSynchronize();
done:
  (void)hipFree(demi);
  (void)hipFree(din);
  (void)hipFree(dout);
}
