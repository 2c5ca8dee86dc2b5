// soi_kernels.hip -- MI355X (gfx950) kernels of the SOI FFT hot path and their C-ABI launchers.
//
// Data layout in HBM (one GPU, P virtual ranks, S = k*P segments):
//   input  alpha[q][i]            q < M rows, i < S lanes (the natural-order N-vector)
//   alpha_tilde[s][R]             s < S segments, R < M_hat (segment-major: the reference's
//                                 all-to-all, pfft.c:553-578, is the identity in this layout)
//   output X[s*M + m]             written over alpha (in place, as the reference does)
//
// Stages (reference -> here):
//   Step 1  parallel_filter_subsampling.cpp:134-524  -> k_filter_sdft   (conv + S-point DFT)
//   Step 2  exchange pfft.c:474-583                  -> identity (layout) / VLC codec kernels
//   Step 3  pfft.c:642 DFTI M_hat-point              -> k_fft_single or k_fft_col + k_fft_row
//           pfft.c:649-693 demodulation              -> fused into the last FFT pass's store
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include "soi_common.hpp"
#include "soi_kernels.h"

using namespace soi;

// =============================================================================================
// Tables (init only; FP64 on the GPU)
// =============================================================================================

// Filter taps and phase factors.  The reference window (pfft.c:42-66) is
//   w[theta][b][i] = e^{i pi t} y(t) / mu,  t = theta*d/n - b - i/S + kappa/2,
//   y(t) = sin(pi tau t)/(pi tau t) * exp(-pi^2 t^2 / sigma), y(0) = 1.
// ref = 1: t as the reference forms it in double, ((double)theta/M_hat - (double)j/N +
//          (double)kappa/(2M))*M with j = b*S + i (no contraction), so the table reproduces the
//          reference's window values; ref = 0: t from its exact rational numerator over 2nS.
// ct = 0 (real taps, exact t only): e^{i pi t} = e^{i pi (t+b)} (-1)^b, so
//          w = post[theta][i] * taps[theta][i][b], taps = (-1)^b y(t)/mu real,
//          post = e^{i pi (theta d/n + kappa/2 - i/S)}.
// ct = 1: taps[theta][i][b] = w (complex), post = 1.
// w_full receives w in the reference's logical layout [theta][b][i] (descriptor field w).
__device__ __forceinline__ double window_y(double t, double tau, double sigma)
{
  if (t == 0.0) return 1.0;
  const double PI_D = 3.14159265358979323846;
  // sin(pi tau t) with the product tau*t carried to double-double
  double ph = __dmul_rn(tau, t);
  double pl = fma(tau, t, -ph);
  double sp, cp;
  sincospi(ph, &sp, &cp);
  double num = fma(M_PI * pl, cp, sp);
  double den = __dmul_rn(__dmul_rn(PI_D, tau), t);
  double e = exp(__ddiv_rn(__dmul_rn(__dmul_rn(__dmul_rn(-PI_D, PI_D), t), t), sigma));
  return __dmul_rn(__ddiv_rn(num, den), e);
}

__global__ void k_make_taps(void *taps, double2 *post, double2 *w_full, int S, int n_mu,
                            int d_mu, int B, int B_pad, double tau, double sigma, long N, long M,
                            long Mh, int ref, int ct)
{
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long total = (long)n_mu * S * B_pad;
  if (idx >= total) return;
  int b = (int)(idx % B_pad);
  long rest = idx / B_pad;
  int i = (int)(rest % S);
  int theta = (int)(rest / S);
  const long den = 2L * n_mu * S;
  const int kappa = B - d_mu;
  const double mu = (double)n_mu / (double)d_mu;
  long num0 = 2L * theta * d_mu * S + (long)kappa * n_mu * S - 2L * i * n_mu;  // (t + b) * den
  // post phase e^{i pi num0/den}: num0 reduced modulo 2*den (period 2 in units of pi)
  long r0 = num0 % (2 * den);
  if (r0 < 0) r0 += 2 * den;
  double sp0, cp0;
  sincospi((double)r0 / (double)den, &sp0, &cp0);
  if (b == 0) post[theta * S + i] = ct ? double2{1.0, 0.0} : double2{cp0, sp0};
  double2 wv = double2{0.0, 0.0};
  if (b < B) {
    double t;
    double sp, cp;
    if (ref) {
      long j = (long)b * S + i;
      double a0 = __ddiv_rn((double)theta, (double)Mh);
      double a1 = __ddiv_rn((double)j, (double)N);
      double a2 = __ddiv_rn((double)kappa, (double)(2 * M));
      t = __dmul_rn(__dadd_rn(__dsub_rn(a0, a1), a2), (double)M);
      sincospi(t, &sp, &cp);
    } else {
      long num = num0 - 2L * b * n_mu * S;
      t = (double)num / (double)den;
      long rr = num % (2 * den);
      if (rr < 0) rr += 2 * den;
      sincospi((double)rr / (double)den, &sp, &cp);
    }
    double y = window_y(t, tau, sigma);
    wv = double2{__ddiv_rn(__dmul_rn(cp, y), mu), __ddiv_rn(__dmul_rn(sp, y), mu)};
    if (w_full) w_full[((long)theta * B + b) * S + i] = wv;
    if (!ct) {
      double tap = __ddiv_rn(y, mu);
      ((double *)taps)[idx] = (b & 1) ? -tap : tap;
    }
  } else if (!ct) {
    ((double *)taps)[idx] = 0.0;
  }
  if (ct) ((double2 *)taps)[idx] = wv;
}

// W_inv[m] = e^{-2 pi i kappa m / (2M)} / w_hat(m/M - 1/2)   (pfft.c:68-82)
// w_hat(t) = [erfc(sqrt(sigma)(t - tau/2)) - erfc(sqrt(sigma)(t + tau/2))] / (2 tau)
__global__ void k_make_winv(double2 *winv, long M, int kappa, double tau, double sigma)
{
  long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m >= M) return;
  double t = (double)m / (double)M - 0.5;
  double ss = sqrt(sigma);
  double y = 1.0 / (2.0 * tau) * (erfc(ss * (t - tau / 2)) - erfc(ss * (t + tau / 2)));
  long r = ((long)kappa * m) % (2 * M);  // angle = -pi * r / M
  double sp, cp;
  sincospi((double)r / (double)M, &sp, &cp);
  winv[m] = double2{cp / y, -sp / y};
}

// tw[m] = e^{-2 pi i m / L} for m < n (exact reduction: m/L in [0,1))
__global__ void k_make_twiddles(double2 *tw, long n, long L, long mult)
{
  long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m >= n) return;
  long e = (m * mult) % L;
  double sp, cp;
  sincospi(2.0 * (double)e / (double)L, &sp, &cp);
  tw[m] = double2{cp, -sp};
}

__global__ void k_f64_to_f32(const double2 *in, float2 *out, long n)
{
  long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m < n) out[m] = float2{(float)in[m].x, (float)in[m].y};
}

// =============================================================================================
// Step 3: M_hat-point FFT per segment, demodulation fused into the last pass
// =============================================================================================
// Single pass: `count` whole segments per block (M_hat small enough for LDS).
template <typename C>
__global__ __launch_bounds__(256) void k_fft_single(const C *__restrict__ at, C *__restrict__ X,
                                                    const C *__restrict__ tw,
                                                    const double2 *__restrict__ winv, long L,
                                                    long M, int S, int count, SoiBatch b)
{
  at = batch_ptr(at, b.at);
  X = batch_ptr(X, b.x);
  extern __shared__ __align__(16) unsigned char lds_raw[];
  C *lds = reinterpret_cast<C *>(lds_raw);
  const int tid = threadIdx.x, nt = blockDim.x;
  const int s0 = blockIdx.x * count;
  const int nseq = min(count, S - s0);
  const int qs = (int)L + 1;
  for (int e = tid; e < nseq * L; e += nt) {
    int q = e / (int)L, j = e - q * (int)L;
    lds[q * qs + j] = at[(long)(s0 + q) * L + j];
  }
  __syncthreads();
  lds_fft<32>(lds, (int)L, nseq, qs, 1, tw);
  for (int e = tid; e < nseq * M; e += nt) {
    int q = e / (int)M, m = e - q * (int)M;
    X[(long)(s0 + q) * M + m] = demod<C>(winv[m], lds[q * qs + m]);
  }
}

// Pass 1 (column FFTs): segment viewed as n1 x n2 row-major (R = n2*r1 + r2).  A block takes
// Tc adjacent columns, runs n1-point FFTs down them, multiplies element m1 of column r2 by
// e^{-2 pi i r2 m1 / M_hat} (two-level table: hi[e / B1] * lo[e % B1]) and stores in place.
template <typename C>
__global__ __launch_bounds__(256) void k_fft_col(C *__restrict__ data, const C *__restrict__ tw1,
                                                 const C *__restrict__ thi,
                                                 const C *__restrict__ tlo, int n1, int n2,
                                                 int Tc, int B1, long Mh, int S, SoiBatch b)
{
  data = batch_ptr(data, b.at);
  extern __shared__ __align__(16) unsigned char lds_raw[];
  C *lds = reinterpret_cast<C *>(lds_raw);
  const int tid = threadIdx.x, nt = blockDim.x;
  const int ncb = n2 / Tc;
  const int bid = xcd_remap(blockIdx.x, gridDim.x);
  const int cb = bid % ncb;
  const int s = bid / ncb;
  if (s >= S) return;
  const int c0 = cb * Tc;
  C *seg = data + (long)s * Mh;
  for (int e = tid; e < n1 * Tc; e += nt) {
    int r1 = e / Tc, q = e - r1 * Tc;
    lds[e] = seg[(long)r1 * n2 + c0 + q];
  }
  __syncthreads();
  lds_fft<32>(lds, n1, Tc, 1, Tc, tw1);
  for (int e = tid; e < n1 * Tc; e += nt) {
    int m1 = e / Tc, q = e - m1 * Tc;
    long ex = (long)(c0 + q) * m1;  // < M_hat
    C t = cmul(thi[ex / B1], tlo[ex % B1]);
    seg[(long)m1 * n2 + c0 + q] = cmul(lds[e], t);
  }
}

// Pass 2 (row FFTs + demodulation): a block takes Tr adjacent rows m1, runs n2-point FFTs along
// them; element m2 of row m1 is y[m1 + n1*m2].  For m < M it stores W_inv[m] * y[m] to X[s*M+m]
// (pfft.c:649-693), Tr adjacent m per contiguous run.
template <typename C>
__global__ __launch_bounds__(256) void k_fft_row(const C *__restrict__ data, C *__restrict__ X,
                                                 const C *__restrict__ tw2,
                                                 const double2 *__restrict__ winv, int n1,
                                                 int n2, int Tr, long Mh, long M, int S,
                                                 SoiBatch b)
{
  data = batch_ptr(data, b.at);
  X = batch_ptr(X, b.x);
  extern __shared__ __align__(16) unsigned char lds_raw[];
  C *lds = reinterpret_cast<C *>(lds_raw);
  const int tid = threadIdx.x, nt = blockDim.x;
  const int nrb = n1 / Tr;
  const int bid = xcd_remap(blockIdx.x, gridDim.x);
  const int s = bid % S;          // segments of one row block are adjacent: W_inv rows shared
  const int rb = bid / S;
  if (rb >= nrb) return;
  const int a = rb * Tr;
  const C *seg = data + (long)s * Mh;
  for (int e = tid; e < n2 * Tr; e += nt) {
    int q = e / n2, r2 = e - q * n2;
    lds[r2 * Tr + q] = seg[(long)(a + q) * n2 + r2];
  }
  __syncthreads();
  lds_fft<32>(lds, n2, Tr, 1, Tr, tw2);
  C *out = X + (long)s * M;
  for (int e = tid; e < n2 * Tr; e += nt) {
    int m2 = e / Tr, q = e - m2 * Tr;
    long m = (long)(a + q) + (long)n1 * m2;
    if (m < M) out[m] = demod<C>(winv[m], lds[e]);
  }
}

// =============================================================================================
// VLC codec (compress.c).  Every 16-element chunk of a stream contributes one field per lane;
// field widths are fixed per stream, so each 32-bit word of lane j is computable directly.
// =============================================================================================
__device__ __forceinline__ int exp_field(double v)
{
  uint64_t u = (uint64_t)__double_as_longlong(v);
  return (int)((u >> 52) & 0x7FF);
}

// per-chunk max exponent (compress.c:46-93).  One block per chunk.
__global__ void k_max_exponent(const double *__restrict__ x, int len, int *__restrict__ out)
{
  __shared__ int red[1024];
  const double *c = x + (long)blockIdx.x * len;
  int m = INT_MIN;
  for (int e = threadIdx.x; e < len; e += blockDim.x) m = max(m, exp_field(c[e]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = (int)((unsigned)red[0] - 1023u);
}

struct VlcParams {
  int i1Len, i2Len, fbits;  // field widths; fbits = bits per element in a lane
  double e1, e2, e3;        // 2^(21-e_max), 2^(52-e_max), 2^31
  double d1, d2;            // 2^(e_max-21), 2^(e_max-52)
};

__device__ __forceinline__ VlcParams vlc_params(int e_max, int e_max_i)
{
  VlcParams v;
  int dl = e_max - e_max_i;
  v.i1Len = 24 - dl;
  v.i2Len = min(32, 55 - dl);
  v.fbits = (v.i1Len > 0 ? v.i1Len : 0) + (v.i2Len > 0 ? v.i2Len : 0);
  v.e1 = ldexp(1.0, 21 - e_max);
  v.e2 = ldexp(1.0, 52 - e_max);
  v.e3 = ldexp(1.0, 31);
  v.d1 = ldexp(1.0, e_max - 21);
  v.d2 = ldexp(1.0, e_max - 52);
  return v;
}

// The bit pattern of one element: i1Len bits of i1 (when i1Len > 0) followed by i2Len bits of
// i2, as a value of fbits <= 56 bits, MSB first.  vec=1: rint + fmsub (compress.c:230-280);
// vec=0: round() + unfused (compress.c:333-342).
__device__ __forceinline__ uint64_t vlc_encode(double x, const VlcParams &v, bool vec)
{
  double d1, d2;
  if (vec) {
    d1 = rint(x * v.e1);
    d2 = rint(fma(x, v.e2, -(d1 * v.e3)));
  } else {
    d1 = round(x * v.e1);
    double pr = __dmul_rn(x, v.e2);
    d2 = round(__dsub_rn(pr, __dmul_rn(d1, v.e3)));
  }
  uint32_t i1 = (uint32_t)(int32_t)d1, i2 = (uint32_t)(int32_t)d2;
  uint64_t f2 = v.i2Len >= 32 ? (uint64_t)i2 : ((uint64_t)i2 & ((1ull << v.i2Len) - 1));
  if (v.i1Len > 0) {
    uint64_t f1 = (uint64_t)i1 & ((1ull << v.i1Len) - 1);
    return (f1 << v.i2Len) | f2;
  }
  return f2;
}

// bits [pos, pos+32) of a lane stream made of consecutive fbits-wide fields (MSB first).
template <typename F>
__device__ __forceinline__ uint32_t vlc_word(long pos, int fbits, long nfields, F field)
{
  uint32_t word = 0;
  long f0, f1;
  if (pos + 31 <= (long)UINT32_MAX) {  // 32-bit divisions: a fraction of the 64-bit sequence
    f0 = (uint32_t)pos / (uint32_t)fbits;
    f1 = (uint32_t)(pos + 31) / (uint32_t)fbits;
  } else {
    f0 = pos / fbits;
    f1 = (pos + 31) / fbits;
  }
  for (long f = f0; f <= f1 && f < nfields; ++f) {
    uint64_t val = field(f);
    long fstart = f * fbits;
    // overlap of [fstart, fstart+fbits) with [pos, pos+32)
    long lo = max(fstart, pos), hi = min(fstart + fbits, pos + 32);
    if (lo >= hi) continue;
    int nb = (int)(hi - lo);
    int shift_out = (int)(fstart + fbits - hi);  // bits of the field after hi
    uint64_t part = (val >> shift_out) & ((1ull << nb) - 1);
    word |= (uint32_t)(part << (pos + 32 - hi));
  }
  return word;
}

// Word wi of a chunk's stream (compress.c:205-372): vector rows (16 lanes interleaved) then the
// scalar tail.
__device__ __forceinline__ uint32_t compress_word(const double *__restrict__ in, int len,
                                                  const VlcParams &v, long wi)
{
  const long nvf = len / 16;                 // fields per lane in the vector body
  const long vrows = (nvf * v.fbits + 31) / 32;
  if (wi < vrows * 16) {
    const long row = wi / 16;
    const int lane = (int)(wi % 16);
    return vlc_word(row * 32, v.fbits, nvf,
                    [&](long f) { return vlc_encode(in[f * 16 + lane], v, true); });
  }
  const long ntail = len - nvf * 16;
  const double *tin = in + nvf * 16;
  return vlc_word((wi - vrows * 16) * 32, v.fbits, ntail,
                  [&](long f) { return vlc_encode(tin[f], v, false); });
}

// Words of a chunk's stream (compress()'s return value)
__device__ __forceinline__ long compress_words(int len, const VlcParams &v)
{
  if (v.i2Len <= 0) return 0;
  const long nvf = len / 16;
  return (nvf * v.fbits + 31) / 32 * 16 + ((len - nvf * 16) * v.fbits + 31) / 32;
}

// Work units of the flat compress: one unit per run of VLC_RUN vector rows of one lane, then one
// per tail word
constexpr int VLC_RUN = 8;
__device__ __forceinline__ long compress_units(int len, const VlcParams &v)
{
  if (v.i2Len <= 0) return 0;
  const long nvf = len / 16;
  const long vrows = (nvf * v.fbits + 31) / 32;
  return (vrows + VLC_RUN - 1) / VLC_RUN * 16 + ((len - nvf * 16) * v.fbits + 31) / 32;
}

// Rows r0 <= r < r1 of lane `lane`'s vector stream (word r*16 + lane), packed in one pass: the
// fields overlapping the rows are encoded once each and shifted in MSB first; fields past the
// lane's nvf are zero padding.  Same words as compress_word.
__device__ __forceinline__ void compress_run(const double *__restrict__ in, int *__restrict__ o,
                                             long nvf, const VlcParams &v, long r0, long r1,
                                             int lane)
{
  const int fbits = v.fbits;
  const long pos = r0 * 32;
  long f = (pos <= (long)UINT32_MAX) ? (long)((uint32_t)pos / (uint32_t)fbits) : pos / fbits;
  int nb = fbits - (int)(pos - f * fbits);  // bits of field f at or after pos
  uint64_t acc = 0;
  int nacc = 0;
  long r = r0;
  while (r < r1) {
    const uint64_t val = f < nvf ? vlc_encode(in[f * 16 + lane], v, true) : 0;
    while (nb > 0 && r < r1) {
      const int take = min(nb, 32 - nacc);
      acc = (acc << take) | ((val >> (nb - take)) & ((1ull << take) - 1));
      nacc += take;
      nb -= take;
      if (nacc == 32) {
        o[r * 16 + lane] = (int)(uint32_t)acc;
        ++r;
        acc = 0;
        nacc = 0;
      }
    }
    ++f;
    nb = fbits;
  }
}

// compress: one thread per output word, blockIdx.y = chunk.  nwords[c] = compress()'s return
// value.
__global__ void k_compress(int *__restrict__ out, long stride, const double *__restrict__ x,
                           int len, int nchunk, int e_max, const int *__restrict__ emi,
                           int *__restrict__ nwords, const int *__restrict__ e_max_dev)
{
  const int c = blockIdx.y;
  if (c >= nchunk) return;
  VlcParams v = vlc_params(e_max_dev ? *e_max_dev : e_max, emi[c]);
  const double *in = x + (long)c * len;
  int *o = out + (long)c * stride;
  const long total = compress_words(len, v);
  if (blockIdx.x == 0 && threadIdx.x == 0) nwords[c] = (int)total;
  for (long wi = blockIdx.x * (long)blockDim.x + threadIdx.x; wi < total;
       wi += (long)gridDim.x * blockDim.x)
    o[wi] = (int)compress_word(in, len, v, wi);
}

// compress over all chunks as one range of work units (upre: exclusive prefix of the chunks'
// unit counts, nchunk + 1 entries, from k_vlc_plan): every block has the same share of units, a
// contiguous span, whatever the chunks' field widths.  A thread finds the chunk of its first unit
// by bisection and then walks forward: chunks hold thousands of units, so the walk rarely moves.
__global__ void k_compress_flat(int *__restrict__ out, long stride, const double *__restrict__ x,
                                int len, int nchunk, const int *__restrict__ e_max_dev,
                                const int *__restrict__ emi, const long *__restrict__ upre)
{
  const long total = upre[nchunk];
  const long span = (total + gridDim.x - 1) / gridDim.x;
  const long u1 = min(total, (blockIdx.x + 1) * span);
  long u = blockIdx.x * span + threadIdx.x;
  if (u >= u1) return;
  int c = 0, hi = nchunk - 1;  // the last chunk c with upre[c] <= u: the one holding unit u
  while (c < hi) {
    const int mid = (c + hi + 1) >> 1;
    if (upre[mid] <= u) c = mid;
    else hi = mid - 1;
  }
  const int e_max = *e_max_dev;
  const long nvf = len / 16;
  long first = upre[c], next = upre[c + 1];
  VlcParams v = vlc_params(e_max, emi[c]);
  long vrows = (nvf * v.fbits + 31) / 32, nvu = (vrows + VLC_RUN - 1) / VLC_RUN * 16;
  for (; u < u1; u += blockDim.x) {
    if (next <= u) {
      do {
        ++c;
        next = upre[c + 1];
      } while (next <= u);
      first = upre[c];
      v = vlc_params(e_max, emi[c]);
      vrows = (nvf * v.fbits + 31) / 32;
      nvu = (vrows + VLC_RUN - 1) / VLC_RUN * 16;
    }
    const long q = u - first;
    const double *in = x + (long)c * len;
    int *o = out + (long)c * stride;
    if (q < nvu) {
      const long r0 = (q >> 4) * VLC_RUN;
      compress_run(in, o, nvf, v, r0, min(vrows, r0 + VLC_RUN), (int)(q & 15));
    } else {
      const long wi = vrows * 16 + (q - nvu);
      o[wi] = (int)compress_word(in, len, v, wi);
    }
  }
}

// extract the fbits-wide field f of a lane stream whose word w is at base[w*wstride]
__device__ __forceinline__ uint64_t vlc_field(const int *base, long wstride, long f, int fbits)
{
  long pos = f * fbits;
  long w0 = pos / 32;
  int off = (int)(pos - w0 * 32);
  uint64_t hi = (uint32_t)base[w0 * wstride];
  uint64_t lo = (off + fbits > 32) ? (uint32_t)base[(w0 + 1) * wstride] : 0u;
  uint64_t bits = (hi << 32) | lo;  // 64-bit window starting at word w0
  int sh = 64 - off - fbits;
  if (off + fbits > 64) {
    // field spans three words (fbits <= 56 and off <= 31 can reach 87 bits)
    uint64_t lo2 = (uint32_t)base[(w0 + 2) * wstride];
    int extra = off + fbits - 64;
    uint64_t top = bits & ((fbits - extra >= 64) ? ~0ull : ((1ull << (64 - off)) - 1));
    return (top << extra) | (lo2 >> (32 - extra));
  }
  return (bits >> sh) & ((fbits >= 64) ? ~0ull : ((1ull << fbits) - 1));
}

__device__ __forceinline__ double vlc_decode(uint64_t fv, const VlcParams &v, bool vec)
{
  int32_t i1 = 0, i2;
  uint64_t f2 = v.i2Len >= 32 ? (fv & 0xFFFFFFFFull) : (fv & ((1ull << v.i2Len) - 1));
  // sign-extend i2Len bits
  i2 = (int32_t)((int64_t)(f2 << (64 - v.i2Len)) >> (64 - v.i2Len));
  if (v.i1Len > 0) {
    uint64_t f1 = (fv >> v.i2Len) & ((1ull << v.i1Len) - 1);
    i1 = (int32_t)((int64_t)(f1 << (64 - v.i1Len)) >> (64 - v.i1Len));
  }
  if (vec) return fma((double)i1, v.d1, (double)i2 * v.d2);       // compress.c:472-483
  return __dadd_rn(__dmul_rn((double)i1, v.d1), __dmul_rn((double)i2, v.d2));  // :529-531
}

// Element e of a chunk decoded from its stream w (compress.c:428-544)
__device__ __forceinline__ double decompress_elem(const int *__restrict__ w, int len,
                                                  const VlcParams &v, long e)
{
  if (v.i2Len <= 0) return 0.0;
  const long nvf = len / 16;
  if (e < nvf * 16)
    return vlc_decode(vlc_field(w + e % 16, 16, e / 16, v.fbits), v, true);
  const long vrows = (nvf * v.fbits + 31) / 32;
  return vlc_decode(vlc_field(w + vrows * 16, 1, e - nvf * 16, v.fbits), v, false);
}

__global__ void k_decompress(double *__restrict__ out, const int *__restrict__ in, long stride,
                             int len, int nchunk, int e_max, const int *__restrict__ emi,
                             const int *__restrict__ e_max_dev)
{
  const int c = blockIdx.y;
  if (c >= nchunk) return;
  VlcParams v = vlc_params(e_max_dev ? *e_max_dev : e_max, emi[c]);
  double *o = out + (long)c * len;
  const int *w = in + (long)c * stride;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < len;
       e += (long)gridDim.x * blockDim.x)
    o[e] = decompress_elem(w, len, v, e);
}

// decompress over all chunks as one range of elements, a contiguous span per block
__global__ void k_decompress_flat(double *__restrict__ out, const int *__restrict__ in,
                                  long stride, int len, int nchunk,
                                  const int *__restrict__ e_max_dev, const int *__restrict__ emi)
{
  const long total = (long)nchunk * len;
  const long span = (total + gridDim.x - 1) / gridDim.x;
  const long g1 = min(total, (blockIdx.x + 1) * span);
  long g = blockIdx.x * span + threadIdx.x;
  if (g >= g1) return;
  const int e_max = *e_max_dev;
  int c = (int)(g / len);
  long e = g - (long)c * len;
  VlcParams v = vlc_params(e_max, emi[c]);
  for (; g < g1; g += blockDim.x, e += blockDim.x) {
    if (e >= len) {
      do {
        e -= len;
        ++c;
      } while (e >= len);
      v = vlc_params(e_max, emi[c]);
    }
    out[g] = decompress_elem(in + (long)c * stride, len, v, e);
  }
}

// =============================================================================================
// Synthetic inputs (seeded, counter based) for tests and the benchmark
// =============================================================================================
__device__ __forceinline__ uint64_t splitmix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t seed, uint64_t n)
{
  return (double)(splitmix64(seed ^ (n * 0xD1B54A32D192ED03ull)) >> 11) * 0x1.0p-53;
}

// uniform[-1,1] re/im (identical to the oracle's orc_uniform_input)
__global__ void k_gen_uniform(double2 *x, long N, uint64_t seed)
{
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= N) return;
  x[n] = double2{2.0 * u01(seed, 2 * n) - 1.0, 2.0 * u01(seed, 2 * n + 1) - 1.0};
}

// sum of tones a_j e^{2 pi i f_j n / N + i phi_j} plus complex Gaussian noise (Box-Muller)
__global__ void k_gen_tones(double2 *x, long N, const double *__restrict__ tf,
                            const double *__restrict__ tph, int ntones, double amp,
                            double noise_sigma, uint64_t seed)
{
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= N) return;
  double re = 0, im = 0;
  for (int j = 0; j < ntones; ++j) {
    // phase in turns: f*n/N exactly reduced using the integer part of f
    double f = tf[j];
    double fi = floor(f), ff = f - fi;
    long fn = ((long)fi % N) * (n % N) % N;  // (fi*n) mod N, exact for N < 2^31
    double turns = (double)fn / (double)N + ff * ((double)n / (double)N) + tph[j];
    turns -= floor(turns);
    double sp, cp;
    sincospi(2.0 * turns, &sp, &cp);
    re += amp * cp;
    im += amp * sp;
  }
  if (noise_sigma > 0) {
    double u1 = u01(seed, 2 * n), u2 = u01(seed, 2 * n + 1);
    double rr = sqrt(-2.0 * log(1.0 - u1)) * noise_sigma;
    double sp, cp;
    sincospi(2.0 * u2, &sp, &cp);
    re += rr * cp;
    im += rr * sp;
  }
  x[n] = double2{re, im};
}

// =============================================================================================
// C-ABI launchers
// =============================================================================================
static inline unsigned nblk(long n, int t) { return (unsigned)((n + t - 1) / t); }

extern "C" int soik_make_tables(void *taps, double2 *post, double2 *w_full, double2 *winv,
                                int S, int n_mu, int d_mu, int B, int B_pad, double tau,
                                double sigma, long N, long M, long Mh, int ref, int ct,
                                hipStream_t st)
{
  long nt = (long)n_mu * S * B_pad;
  hipLaunchKernelGGL(k_make_taps, dim3(nblk(nt, 256)), dim3(256), 0, st, taps, post, w_full, S,
                     n_mu, d_mu, B, B_pad, tau, sigma, N, M, Mh, ref, ct);
  hipLaunchKernelGGL(k_make_winv, dim3(nblk(M, 256)), dim3(256), 0, st, winv, M, B - d_mu, tau,
                     sigma);
  return (int)hipGetLastError();
}

extern "C" int soik_make_twiddles(double2 *tw, long n, long L, long mult, hipStream_t st)
{
  hipLaunchKernelGGL(k_make_twiddles, dim3(nblk(n, 256)), dim3(256), 0, st, tw, n, L, mult);
  return (int)hipGetLastError();
}

static void set_lds_attr(const void *f)
{
  (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

template <typename C>
static int fft_single(const C *at, C *X, const C *tw, const double2 *winv, long L, long M, int S,
                      int count, SoiBatch b, hipStream_t st)
{
  static bool once = false;
  if (!once) { set_lds_attr((const void *)k_fft_single<C>); once = true; }
  size_t lds = (size_t)count * (L + 1) * sizeof(C);
  hipLaunchKernelGGL(k_fft_single<C>, dim3((S + count - 1) / count, b.n > 0 ? b.n : 1), dim3(256),
                     lds, st, at, X, tw, winv, L, M, S, count, b);
  return (int)hipGetLastError();
}

template <typename C>
static int fft_col(C *data, const C *tw1, const C *thi, const C *tlo, int n1, int n2, int Tc,
                   int B1, long Mh, int S, SoiBatch b, hipStream_t st)
{
  static bool once = false;
  if (!once) { set_lds_attr((const void *)k_fft_col<C>); once = true; }
  size_t lds = (size_t)n1 * Tc * sizeof(C);
  unsigned grid = (unsigned)((n2 / Tc) * S);
  hipLaunchKernelGGL(k_fft_col<C>, dim3(grid, b.n > 0 ? b.n : 1), dim3(256), lds, st, data, tw1,
                     thi, tlo, n1, n2, Tc, B1, Mh, S, b);
  return (int)hipGetLastError();
}

template <typename C>
static int fft_row(const C *data, C *X, const C *tw2, const double2 *winv, int n1, int n2,
                   int Tr, long Mh, long M, int S, SoiBatch b, hipStream_t st)
{
  static bool once = false;
  if (!once) { set_lds_attr((const void *)k_fft_row<C>); once = true; }
  size_t lds = (size_t)n2 * Tr * sizeof(C);
  unsigned grid = (unsigned)((n1 / Tr) * S);
  hipLaunchKernelGGL(k_fft_row<C>, dim3(grid, b.n > 0 ? b.n : 1), dim3(256), lds, st, data, X,
                     tw2, winv, n1, n2, Tr, Mh, M, S, b);
  return (int)hipGetLastError();
}

extern "C" int soik_fft_single(const double2 *at, double2 *X, const double2 *tw,
                               const double2 *winv, long L, long M, int S, int count,
                               SoiBatch b, hipStream_t st)
{
  return fft_single<double2>(at, X, tw, winv, L, M, S, count, b, st);
}

extern "C" int soik_fft_col(double2 *data, const double2 *tw1, const double2 *thi,
                            const double2 *tlo, int n1, int n2, int Tc, int B1, long Mh, int S,
                            SoiBatch b, hipStream_t st)
{
  return fft_col<double2>(data, tw1, thi, tlo, n1, n2, Tc, B1, Mh, S, b, st);
}

extern "C" int soik_fft_row(const double2 *data, double2 *X, const double2 *tw2,
                            const double2 *winv, int n1, int n2, int Tr, long Mh, long M, int S,
                            SoiBatch b, hipStream_t st)
{
  return fft_row<double2>(data, X, tw2, winv, n1, n2, Tr, Mh, M, S, b, st);
}

// FP32 form: complex64 data and tables; W_inv stays FP64
extern "C" int soik_fft_single_f32(const float2 *at, float2 *X, const float2 *tw,
                                   const double2 *winv, long L, long M, int S, int count,
                                   SoiBatch b, hipStream_t st)
{
  return fft_single<float2>(at, X, tw, winv, L, M, S, count, b, st);
}

extern "C" int soik_fft_col_f32(float2 *data, const float2 *tw1, const float2 *thi,
                                const float2 *tlo, int n1, int n2, int Tc, int B1, long Mh,
                                int S, SoiBatch b, hipStream_t st)
{
  return fft_col<float2>(data, tw1, thi, tlo, n1, n2, Tc, B1, Mh, S, b, st);
}

extern "C" int soik_fft_row_f32(const float2 *data, float2 *X, const float2 *tw2,
                                const double2 *winv, int n1, int n2, int Tr, long Mh, long M,
                                int S, SoiBatch b, hipStream_t st)
{
  return fft_row<float2>(data, X, tw2, winv, n1, n2, Tr, Mh, M, S, b, st);
}

extern "C" int soik_max_exponent(const double *x, int len, int nchunk, int *out, hipStream_t st)
{
  if (nchunk <= 0) return 0;
  hipLaunchKernelGGL(k_max_exponent, dim3(nchunk), dim3(256), 0, st, x, len, out);
  return (int)hipGetLastError();
}

extern "C" int soik_compress(int *out, long stride, const double *x, int len, int nchunk,
                             int e_max, const int *emi, int *nwords, const int *e_max_dev,
                             hipStream_t st)
{
  if (nchunk <= 0) return 0;
  long maxw = (long)len * 56 / 32 / 1 + 64;
  // blocks per chunk: enough for about 2048 blocks in all (the kernel strides over the chunk's
  // words), not one thread per possible word: a chunk of short fields has few words, and a
  // grid sized for 55-bit fields would be mostly idle blocks
  unsigned gx = nblk(maxw, 256);
  const unsigned cap = (unsigned)((2048 + nchunk - 1) / nchunk);
  if (gx > cap) gx = cap;
  hipLaunchKernelGGL(k_compress, dim3(gx, nchunk), dim3(256), 0, st, out, stride, x, len, nchunk,
                     e_max, emi, nwords, e_max_dev);
  return (int)hipGetLastError();
}

extern "C" int soik_decompress(double *out, const int *in, long stride, int len, int nchunk,
                               int e_max, const int *emi, const int *e_max_dev, hipStream_t st)
{
  if (nchunk <= 0) return 0;
  unsigned gx = nblk(len, 256);
  const unsigned cap = (unsigned)((2048 + nchunk - 1) / nchunk);  // about 2048 blocks in all
  if (gx > cap) gx = cap;
  hipLaunchKernelGGL(k_decompress, dim3(gx, nchunk), dim3(256), 0, st, out, in, stride, len,
                     nchunk, e_max, emi, e_max_dev);
  return (int)hipGetLastError();
}

// VLC planning on the device (pfft.c:316-343, 390-440): from the per-chunk exponents of the
// S x P chunks (chunk (s, p) = s*P + p), the per-segment maximum (the reference's
// MPI_Allreduce(MAX) over ranks), the global maximum, every chunk's segment exponent, the
// per-segment send counts in doubles and the cost-balanced segment boundaries.  One block; the
// boundary walk is sequential in thread 0 with the reference's double arithmetic, unfused, so
// the boundaries are bit-identical to the host restatement.
// out: [0] global maximum, [1 .. S] segment maxima, [1+S .. 2S] send counts,
//      [1+2S .. 1+2S+P] segment boundaries; pcs: S + 1 doubles of scratch.
__global__ void k_vlc_plan(const int *__restrict__ exps, int S, int P, int k, long l, long Mh,
                           double ratio, int *__restrict__ emi, int *__restrict__ out,
                           double *__restrict__ pcs_g, long *__restrict__ upre,
                           int *__restrict__ nwords)
{
#pragma clang fp contract(off)
  // segment maxima and prefix costs in LDS when S is small enough (the serial walk below then
  // reads no global memory), else in the caller's scratch
  constexpr int SMAX = 2048;
  __shared__ int smax, segs[SMAX];
  __shared__ double pcs_s[SMAX + 1];
  const bool inl = S <= SMAX;
  double *pcs = inl ? pcs_s : pcs_g;
  if (threadIdx.x == 0) smax = INT_MIN;
  __syncthreads();
  for (int s = threadIdx.x; s < S; s += blockDim.x) {
    int m = INT_MIN;
    for (int p = 0; p < P; ++p) m = max(m, exps[s * P + p]);
    out[1 + s] = m;
    if (inl) segs[s] = m;
    atomicMax(&smax, m);
  }
  __syncthreads();
  const int total = smax;
  for (int c = threadIdx.x; c < S * P; c += blockDim.x) emi[c] = inl ? segs[c / P] : out[1 + c / P];
  if (upre) {
    // work units per chunk (all chunks of a segment alike) and their exclusive prefix over the
    // chunks, for k_compress_flat; nwords = compress()'s return value per chunk
    const int len = (int)(2 * l);
    __shared__ long useg_total;
    if (threadIdx.x == 0) {
      long acc = 0;
      for (int s = 0; s < S; ++s) {
        if (inl) pcs_s[s] = (double)acc;  // reused below as each segment's first unit (exact)
        acc += compress_units(len, vlc_params(total, inl ? segs[s] : out[1 + s])) * P;
      }
      useg_total = acc;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < S * P; c += blockDim.x) {
      const int s = c / P, p = c % P;
      const VlcParams v = vlc_params(total, inl ? segs[s] : out[1 + s]);
      long first = 0;
      if (inl) {
        first = (long)pcs_s[s];
      } else {
        for (int q = 0; q < s; ++q)
          first += compress_units(len, vlc_params(total, out[1 + q])) * P;
      }
      upre[c] = first + compress_units(len, v) * p;
      nwords[c] = (int)compress_words(len, v);
    }
    if (threadIdx.x == 0) upre[S * P] = useg_total;
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  out[0] = total;
  pcs[0] = 0;
  for (int p = 0; p < P; ++p)
    for (int ik = 0; ik < k; ++ik) {
      const int s = p * k + ik;
      int bits = 52 + 3 - (total - (inl ? segs[s] : out[1 + s]));
      if (bits > 31) ++bits;
      const int lenPerSimdLane = (int)((l * 2 / 16 * bits + 31) / 32);
      const int remainder = (int)((l * 2 % 16 * bits + 31) / 32);
      int sendLen = lenPerSimdLane * 16 + remainder;
      if (bits <= 0) sendLen = 0;
      const int cnt = (sendLen + 1) / 2;
      out[1 + S + s] = cnt;
      pcs[s + 1] = pcs[s] + ratio * cnt / 2 / Mh + 1;
    }
  int *bd = out + 1 + 2 * S;
  bd[0] = 0;
  bd[P] = S;
  const double costPerRank = pcs[S] / P;
  int s = 1;
  for (int p = 1; p < P;) {
    const double target = p * costPerRank;
    if (pcs[s] >= target) {
      bd[p] = (pcs[s] - target > target - pcs[s - 1]) ? s - 1 : s;
      ++p;
    } else {
      ++s;
    }
  }
}

extern "C" int soik_vlc_plan(const int *exps, int S, int P, int k, long l, long Mh, double ratio,
                             int *emi, int *out, double *pcs, long *upre, int *nwords,
                             hipStream_t st)
{
  hipLaunchKernelGGL(k_vlc_plan, dim3(1), dim3(256), 0, st, exps, S, P, k, l, Mh, ratio, emi, out,
                     pcs, upre, nwords);
  return (int)hipGetLastError();
}

// Flat compress / decompress of the nchunk chunks of one VLC exchange (after soik_vlc_plan)
extern "C" int soik_compress_flat(int *out, long stride, const double *x, int len, int nchunk,
                                  const int *e_max_dev, const int *emi, const long *upre,
                                  hipStream_t st)
{
  hipLaunchKernelGGL(k_compress_flat, dim3(2048), dim3(256), 0, st, out, stride, x, len, nchunk,
                     e_max_dev, emi, upre);
  return (int)hipGetLastError();
}

extern "C" int soik_decompress_flat(double *out, const int *in, long stride, int len, int nchunk,
                                    const int *e_max_dev, const int *emi, hipStream_t st)
{
  hipLaunchKernelGGL(k_decompress_flat, dim3(2048), dim3(256), 0, st, out, in, stride, len,
                     nchunk, e_max_dev, emi);
  return (int)hipGetLastError();
}

extern "C" int soik_gen_uniform(double2 *x, long N, uint64_t seed, hipStream_t st)
{
  hipLaunchKernelGGL(k_gen_uniform, dim3(nblk(N, 256)), dim3(256), 0, st, x, N, seed);
  return (int)hipGetLastError();
}

extern "C" int soik_gen_tones(double2 *x, long N, const double *tf, const double *tph, int ntones,
                              double amp, double noise_sigma, uint64_t seed, hipStream_t st)
{
  hipLaunchKernelGGL(k_gen_tones, dim3(nblk(N, 256)), dim3(256), 0, st, x, N, tf, tph, ntones,
                     amp, noise_sigma, seed);
  return (int)hipGetLastError();
}

extern "C" int soik_f64_to_f32(const double2 *in, float2 *out, long n, hipStream_t st)
{
  hipLaunchKernelGGL(k_f64_to_f32, dim3(nblk(n, 256)), dim3(256), 0, st, in, out, n);
  return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Allocation-free compress (soi_compress_noalloc in compress.h): the per-chunk segment exponents
// and word counts that compress() stages through hipMalloc'ed buffers live in a file-scope
// device array instead.  Its device address comes from hipGetSymbolAddress, once per device on
// the first call there (the array's name is a host-side stand-in in host code and is never
// passed to a kernel or a copy).  Every call on a device shares that device's array: calls
// that overlap on one device race.  The kernel writes every word count it reports, and the
// exponents are copied in before the kernel runs, so nothing needs clearing between calls.
// ---------------------------------------------------------------------------------------------
#include <mutex>

#define SOI_NOALLOC_MAX_CHUNKS 4096
#define SOI_NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) int scratch[2 * SOI_NOALLOC_MAX_CHUNKS];
static int *g_scratch_ptr[SOI_NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

extern "C" void soi_internal_set_error(int status, const char *msg);

static int *scratch_for_current_device(void)
{
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SOI_NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<int *>(p);
  }
  return g_scratch_ptr[dev];
}

extern "C" int soi_compress_noalloc(int *out, long stride_words, const double *x, int len,
                                    int nchunk, int e_max, const int *e_max_i, int *nwords,
                                    void *stream)
{
  if (!out || !x || !e_max_i || !nwords || len < 0 || nchunk < 1 ||
      stride_words < (long)len * 56 / 32 + 32) {
    soi_internal_set_error(1, "soi_compress_noalloc: needs device out and x, host e_max_i and "
                              "nwords, nchunk >= 1, stride_words >= len*56/32 + 32");
    return 1;
  }
  if (nchunk > SOI_NOALLOC_MAX_CHUNKS) {
    soi_internal_set_error(1, "soi_compress_noalloc: nchunk exceeds the scratch maximum of 4096 "
                              "chunks");
    return 1;
  }
  int *ptr = scratch_for_current_device();
  if (!ptr) {
    soi_internal_set_error(5, "soi_compress_noalloc: hipGetSymbolAddress of the scratch failed");
    return 5;
  }
  hipStream_t st = (hipStream_t)stream;
  int *emi = ptr, *nw = ptr + SOI_NOALLOC_MAX_CHUNKS;
  int e = (int)hipMemcpyAsync(emi, e_max_i, sizeof(int) * nchunk, hipMemcpyHostToDevice, st);
  if (!e) e = soik_compress(out, stride_words, x, len, nchunk, e_max, emi, nw, nullptr, st);
  if (!e) e = (int)hipMemcpyAsync(nwords, nw, sizeof(int) * nchunk, hipMemcpyDeviceToHost, st);
  if (!e) e = (int)hipStreamSynchronize(st);
  if (e) {
    soi_internal_set_error(5, hipGetErrorString((hipError_t)e));
    return 5;
  }
  soi_internal_set_error(0, "");
  return 0;
}
