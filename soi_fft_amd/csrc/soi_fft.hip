// soi_fft.hip -- M_hat-point segment FFTs of the SOI FFT on MI355X (reference: the DFTI transform
// pfft.c:227-228 runs at pfft.c:642, then the demodulation pfft.c:649-693).
//
// Four-step split M_hat = n1 * n2, R = n2*r1 + r2, m = m1 + n1*m2:
//   pass 1 (columns): for each column r2, Z[r2][m1] = sum_r1 x[n2 r1 + r2] w_n1^{r1 m1}, times the
//                     twiddle w_Mhat^{r2 m1}, written back in place at n2*m1 + r2;
//   pass 2 (rows):    y[m1 + n1 m2] = sum_r2 Z[r2][m1] w_n2^{r2 m2}, demodulated and stored as
//                     X[s*M + m] = W_inv[m] * y[m] for m < M.
// Both passes keep their transforms in registers (E values per thread, Stockham stages with the
// radices of the template list) and exchange through LDS between stages.  A block takes T = 8
// adjacent columns (pass 1) or rows (pass 2) so every global access moves 128-byte runs; blocks
// are remapped so neighbouring tiles share an XCD (and its L2).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "soi_common.hpp"
#include "soi_kernels.h"

using namespace soi;

// ---------------------------------------------------------------------------------------------
// radix-10 and radix-16 forward DFTs in registers
// ---------------------------------------------------------------------------------------------
template <typename C> __device__ __forceinline__ void dft16(C *a)
{
  typedef decltype(a[0].x) T;
  // 16 = 4 x 4: columns n = n1 + 4 n2, first DFT4 over n2, twiddle w16^{n1 k2}, DFT4 over n1
  C b[16];
#pragma unroll
  for (int n1 = 0; n1 < 4; ++n1) {
    C x0 = a[n1], x1 = a[n1 + 4], x2 = a[n1 + 8], x3 = a[n1 + 12];
    dft4(x0, x1, x2, x3);
    b[n1 * 4 + 0] = x0;
    b[n1 * 4 + 1] = x1;
    b[n1 * 4 + 2] = x2;
    b[n1 * 4 + 3] = x3;
  }
  const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173;
  const T r2 = (T)0.70710678118654752440;
  // twiddles w16^{n1 k2} = e^{-2 pi i n1 k2 / 16}
  // n1 = 1: k2 = 1,2,3 -> w^1, w^2, w^3 ; n1 = 2: w^2, w^4, w^6 ; n1 = 3: w^3, w^6, w^9
  const C w1 = C{c1, -s1}, w2 = C{r2, -r2}, w3 = C{s1, -c1};
  const C w6 = C{-r2, -r2}, w9 = C{-c1, s1};
  b[1 * 4 + 1] = cmul(b[5], w1);
  b[1 * 4 + 2] = cmul(b[6], w2);
  b[1 * 4 + 3] = cmul(b[7], w3);
  b[2 * 4 + 1] = cmul(b[9], w2);
  b[2 * 4 + 2] = cmul_mi(b[10]);
  b[2 * 4 + 3] = cmul(b[11], w6);
  b[3 * 4 + 1] = cmul(b[13], w3);
  b[3 * 4 + 2] = cmul(b[14], w6);
  b[3 * 4 + 3] = cmul(b[15], w9);
#pragma unroll
  for (int k2 = 0; k2 < 4; ++k2) {
    C x0 = b[0 * 4 + k2], x1 = b[1 * 4 + k2], x2 = b[2 * 4 + k2], x3 = b[3 * 4 + k2];
    dft4(x0, x1, x2, x3);
    // output k = k2 + 4 k1
    a[k2 + 0] = x0;
    a[k2 + 4] = x1;
    a[k2 + 8] = x2;
    a[k2 + 12] = x3;
  }
}

template <typename C> __device__ __forceinline__ void dft10(C *a)
{
  typedef decltype(a[0].x) T;
  // 10 = 2 x 5 (Good-Thomas free form): n = n1 + 2 n2 ... use n = 5 n1 + n2? Use Cooley-Tukey:
  // n = n1 + 2 n2 (n1 < 2, n2 < 5): DFT5 over n2 for each n1, twiddle w10^{n1 k2}, DFT2 over n1.
  C e[5] = {a[0], a[2], a[4], a[6], a[8]};
  C o[5] = {a[1], a[3], a[5], a[7], a[9]};
  dft5(e[0], e[1], e[2], e[3], e[4]);
  dft5(o[0], o[1], o[2], o[3], o[4]);
  const T c1 = (T)0.80901699437494742410, s1 = (T)0.58778525229247312917;  // cos, sin 36 deg
  const T c2 = (T)0.30901699437494742410, s2 = (T)0.95105651629515357212;  // cos, sin 72 deg
  o[1] = cmul(o[1], C{c1, -s1});
  o[2] = cmul(o[2], C{c2, -s2});
  o[3] = cmul(o[3], C{-c2, -s2});
  o[4] = cmul(o[4], C{-c1, -s1});
#pragma unroll
  for (int k2 = 0; k2 < 5; ++k2) {
    a[k2] = cadd(e[k2], o[k2]);
    a[k2 + 5] = csub(e[k2], o[k2]);
  }
}

template <int R, typename C> __device__ __forceinline__ void dft_any(C *a)
{
  if constexpr (R == 16) dft16(a);
  else if constexpr (R == 10) dft10(a);
  else dft_r<R>(a);
}

// ---------------------------------------------------------------------------------------------
// compile-time radix schedules
// ---------------------------------------------------------------------------------------------
template <int... Rs> struct Radices {
  static constexpr int n = sizeof...(Rs);
  static constexpr int r[n] = {Rs...};
  static constexpr int at(int s) { return r[s]; }
  static constexpr int span(int s)  // product of radices before stage s
  {
    int p = 1;
    for (int i = 0; i < s; ++i) p *= r[i];
    return p;
  }
};

// LDS layouts of the T/NQP sequences exchanged at once (one real per slot; real and imaginary
// halves go through in turn).
//   QFast:    sequence index fastest, (pos + pos/16)*TG + q -- suits waves that span the T
//             sequences at a few positions (the column pass, whose lanes run over q first)
//   QFast8:   the same with one pad slot per 8 positions: the 8-wide 2560-point columns, whose
//             radix-10 stage writes positions 10 apart (complex64: 1.63 -> 1.40 ms for that
//             variant; the default 4-wide one keeps QFast)
//   SeqMajor: one padded run per sequence, q*LP + pos + pos/16 with LP = 2 (mod 16) -- suits
//             waves that hold consecutive positions of one sequence (the row pass); the LP offset
//             keeps the T-sequence transpose reads of the row pass on distinct banks
template <int L, int TG> struct QFast {
  static constexpr bool full = false;  // complex128: exchange in real / imaginary rounds
  static constexpr int size = (L + L / 16) * TG;
  static __device__ __forceinline__ int at(int pos, int q) { return (pos + (pos >> 4)) * TG + q; }
};
template <int L, int TG> struct QFast8 {
  static constexpr bool full = false;
  static constexpr int size = (L + L / 8) * TG;
  static __device__ __forceinline__ int at(int pos, int q) { return (pos + (pos >> 3)) * TG + q; }
};
//   QFull:    whole complex values, no padding: L*TG*16 bytes (160 KiB for 2560 points x 4
//             complex128 sequences, the most one block may hold).  Position bits 0-3 are XORed
//             with bits 4-7 (a bijection within each run of 16) to spread the radix-10 stage's
//             writes, 10 positions apart, over the banks.
template <int L, int TG> struct QFull {
  static_assert(L % 16 == 0, "swizzle within runs of 16");
  static constexpr bool full = true;
  static constexpr int size = L * TG;
  static __device__ __forceinline__ int at(int pos, int q)
  {
    return ((pos & ~15) | ((pos ^ (pos >> 4)) & 15)) * TG + q;
  }
};
template <int L, int TG> struct SeqMajor {
  static constexpr bool full = false;
  static constexpr int LP0 = L + L / 16;
  static constexpr int LP = LP0 + ((2 - LP0 % 16) + 16) % 16;
  static constexpr int size = LP * TG;
  static __device__ __forceinline__ int at(int pos, int q) { return q * LP + pos + (pos >> 4); }
};

// Complex64 exchanges move whole complex values in one round when the layout fits in LDS: half
// the LDS instructions and barriers of the real / imaginary rounds.  Complex128 tiles keep the
// rounds in the padded layouts (a padded 2560 x 4 tile exchanged whole needs two groups of
// sequences, measured slower: 1.83 against 1.60 ms at N = 2^27); the unpadded QFull layout sets
// `full` and holds a whole 2560 x 4 complex128 tile in 160 KiB (1.59 -> 1.51 ms).
template <typename LY, typename C> constexpr bool exch_full()
{
  return (sizeof(real_t<C>) == 4 || LY::full) && (size_t)LY::size * sizeof(C) <= 160 * 1024;
}
template <typename LY, typename C> constexpr size_t exch_lds_bytes()
{
  return (size_t)LY::size * (exch_full<LY, C>() ? sizeof(C) : sizeof(real_t<C>));
}

// Exchange a thread's stage outputs through LDS into the next stage's input order.  Outputs of
// stage s (radix R, span Ns) for butterfly j land at (j - j%Ns)*R + j%Ns + k*Ns; stage s+1
// (radix R2) butterfly j2 reads j2 + k*L/R2.  Done in parts: real then imaginary halves (or whole
// complex values, exch_full), and NQP groups of the T sequences, so the LDS footprint is
// L*(T/NQP)*8 bytes (plus padding).
template <int L, int E, int T, int NQP, int TS, int R, int NS, int R2, typename LY, typename C>
__device__ __forceinline__ void exchange(C (&v)[E], real_t<C> *lds, int q, int ts)
{
  constexpr int NB = E / R, NB2 = E / R2, TG = T / NQP;
  const int qg = q / TG, ql = q % TG;
  if constexpr (exch_full<LY, C>()) {
    C *ldc = reinterpret_cast<C *>(lds);
#pragma unroll
    for (int g = 0; g < NQP; ++g) {
      if (qg == g) {
#pragma unroll
        for (int c = 0; c < NB; ++c) {
          const int j = ts + TS * c;
          const int jm = j % NS;
          const int base = (j - jm) * R + jm;
#pragma unroll
          for (int k = 0; k < R; ++k) ldc[LY::at(base + k * NS, ql)] = v[c * R + k];
        }
      }
      lds_barrier();
      if (qg == g) {
#pragma unroll
        for (int c = 0; c < NB2; ++c) {
          const int j = ts + TS * c;
#pragma unroll
          for (int k = 0; k < R2; ++k) v[c * R2 + k] = ldc[LY::at(j + k * (L / R2), ql)];
        }
      }
      lds_barrier();
    }
    return;
  }
#pragma unroll
  for (int g = 0; g < NQP; ++g) {
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      if (qg == g) {
#pragma unroll
        for (int c = 0; c < NB; ++c) {
          const int j = ts + TS * c;
          const int jm = j % NS;
          const int base = (j - jm) * R + jm;
#pragma unroll
          for (int k = 0; k < R; ++k)
            lds[LY::at(base + k * NS, ql)] = half ? v[c * R + k].y : v[c * R + k].x;
        }
      }
      lds_barrier();
      if (qg == g) {
#pragma unroll
        for (int c = 0; c < NB2; ++c) {
          const int j = ts + TS * c;
#pragma unroll
          for (int k = 0; k < R2; ++k) {
            const real_t<C> val = lds[LY::at(j + k * (L / R2), ql)];
            if (half) v[c * R2 + k].y = val; else v[c * R2 + k].x = val;
          }
        }
      }
      lds_barrier();
    }
  }
}

constexpr int bit_len(int x) { return x ? 1 + bit_len(x >> 1) : 0; }

// twiddle + DFT of every butterfly of stage s held in registers.  The R - 1 twiddles w^k of a
// butterfly come from the log2 R table reads w^(2^b): w^k is the product of the powers of k's
// set bits (at most three extra roundings for R = 16), so a thread issues a few loads per
// butterfly instead of R - 1.
template <int L, int E, int TS, int R, int NS, typename C>
__device__ __forceinline__ void stage(C (&v)[E], int ts, const C *__restrict__ tw)
{
  constexpr int NB = E / R;
  constexpr int NBITS = bit_len(R - 1);
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    const int j = ts + TS * c;
    if constexpr (NS > 1) {
      const int jm = j % NS;
      constexpr int step = L / (NS * R);
      C p[NBITS];
#pragma unroll
      for (int b = 0; b < NBITS; ++b) p[b] = tw[(jm * step) << b];
#pragma unroll
      for (int k = 1; k < R; ++k) {
        C w = C{0, 0};
        bool first = true;
#pragma unroll
        for (int b = 0; b < NBITS; ++b)
          if (k & (1 << b)) {
            w = first ? p[b] : cmul(w, p[b]);
            first = false;
          }
        v[c * R + k] = cmul(v[c * R + k], w);
      }
    }
    dft_any<R>(&v[c * R]);
  }
}

// Run stages 0..n-1 on register data already loaded for stage 0.
template <int L, int E, int T, int NQP, int TS, typename RS, typename LY, int S0, typename C>
__device__ __forceinline__ void run_stages(C (&v)[E], real_t<C> *lds, int q, int ts,
                                           const C *__restrict__ tw)
{
  constexpr int R = RS::at(S0);
  constexpr int NS = RS::span(S0);
  stage<L, E, TS, R, NS, C>(v, ts, tw);
  if constexpr (S0 + 1 < RS::n) {
    exchange<L, E, T, NQP, TS, R, NS, RS::at(S0 + 1), LY, C>(v, lds, q, ts);
    run_stages<L, E, T, NQP, TS, RS, LY, S0 + 1, C>(v, lds, q, ts, tw);
  }
}

// After the last stage, value v[c*R + k] of butterfly j = ts + TS*c is output index
// (j - j%NS)*R + j%NS + k*NS with NS = L/R.
template <int L, int TS, typename RS>
__device__ __forceinline__ int out_index(int ts, int c, int k)
{
  constexpr int R = RS::at(RS::n - 1);
  constexpr int NS = L / R;
  const int j = ts + TS * c;
  const int jm = j % NS;
  return (j - jm) * R + jm + k * NS;
}

// ---------------------------------------------------------------------------------------------
// Pass 1: T adjacent columns of length L = n1 (stride n2), twiddled, in place
// ---------------------------------------------------------------------------------------------
// waves per SIMD the register budget is sized for, with BPC blocks resident per CU
constexpr int fft_minw(int BPC, int threads) { return (BPC * threads / 64 + 3) / 4; }

template <int L, int E, int T, int NQP, int BPC, typename RS, typename LY, typename C, bool TILED>
__global__ __launch_bounds__(T * (L / E), fft_minw(BPC, T * (L / E))) void k_fft_cols(C *__restrict__ data,
                                                          const C *__restrict__ tw,
                                                          const C *__restrict__ thi,
                                                          const C *__restrict__ tlo,
                                                          int n2, int lb1, long Mh, int S,
                                                          C *__restrict__ tout, SoiBatch b)
{
  constexpr int TS = L / E;
  data = batch_ptr(data, b.at);
  if constexpr (TILED) tout = batch_ptr(tout, b.tb);
  extern __shared__ __align__(16) unsigned char lds_raw[];
  real_t<C> *ldsd = reinterpret_cast<real_t<C> *>(lds_raw);
  const int tid = threadIdx.x;
  const int q = tid % T, ts = tid / T;
  const int ncb = n2 / T;
  const int bid = xcd_remap(blockIdx.x, gridDim.x);
  const int cb = bid % ncb, s = bid / ncb;
  if (s >= S) return;
  const int c0 = cb * T;
  C *seg = data + (long)s * Mh + c0 + q;
  const int n2s = n2;

  C v[E];
  constexpr int R0 = RS::at(0);
#pragma unroll
  for (int c = 0; c < E / R0; ++c)
#pragma unroll
    for (int k = 0; k < R0; ++k) v[c * R0 + k] = seg[(unsigned)(ts + TS * c + k * (L / R0)) * (unsigned)n2s];

  run_stages<L, E, T, NQP, TS, RS, LY, 0, C>(v, ldsd, q, ts, tw);

  // twiddle w_Mhat^{r2 m1}, exponent below M_hat < 2^32, from the two-level table
  // (hi[e >> lb1] * lo[e & mask]).  Where the register budget allows (complex64, or one block
  // of at most 512 threads per CU) it is factored: m1 = out_index(c, 0) + k*L/RL, so the twiddle is
  // w^{r2 out_index(c, 0)} times w^{r2 k L/RL}, and a thread reads E/RL + RL - 1 table pairs
  // instead of E.
  constexpr int RL = RS::at(RS::n - 1);
  constexpr bool FACT = sizeof(real_t<C>) == 4 || (BPC == 1 && T * (L / E) <= 512);
  const unsigned r2 = (unsigned)(c0 + q), lmask = (1u << lb1) - 1;
  C tk[RL];
  if constexpr (FACT) {
#pragma unroll
    for (int k = 1; k < RL; ++k) {
      const unsigned ex = r2 * (unsigned)(k * (L / RL));
      tk[k] = cmul(thi[ex >> lb1], tlo[ex & lmask]);
    }
  }
#pragma unroll
  for (int c = 0; c < E / RL; ++c) {
    C tc = C{1, 0};
    if constexpr (FACT) {
      const unsigned ex0 = r2 * (unsigned)out_index<L, TS, RS>(ts, c, 0);
      tc = cmul(thi[ex0 >> lb1], tlo[ex0 & lmask]);
    }
#pragma unroll
    for (int k = 0; k < RL; ++k) {
      const int m1 = out_index<L, TS, RS>(ts, c, k);
      C t;
      if constexpr (FACT) {
        t = k ? cmul(tc, tk[k]) : tc;
      } else {
        const unsigned ex = r2 * (unsigned)m1;
        t = cmul(thi[ex >> lb1], tlo[ex & lmask]);
      }
      if constexpr (TILED)  // 8-row tiles: element (m1, r2) at ((m1/8)*n2 + r2)*8 + m1%8
        tout[(long)s * Mh + (long)(((unsigned)(m1 >> 3) * (unsigned)n2 + (unsigned)(c0 + q)) * 8u +
                                   (unsigned)(m1 & 7))] = cmul(v[c * RL + k], t);
      else
        seg[(unsigned)m1 * (unsigned)n2s] = cmul(v[c * RL + k], t);
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Pass 2 over the 8-row tiles written by pass 1 (TILED): the block's 8 rows are one contiguous
// run of 8*L values.  Lanes run over the 8 rows first, so the loads are contiguous and the
// demodulated stores X[a + q + n1*m2] leave in 128-byte (64-byte for complex64) runs with no
// transpose; exchanges use the sequence-index-fastest LDS layout.  Row blocks of one segment
// are adjacent in launch order, so the stores of concurrently running blocks are adjacent too.
// ---------------------------------------------------------------------------------------------
template <int L, int E, int BPC, typename RS, typename C>
__global__ __launch_bounds__(8 * (L / E), fft_minw(BPC, 8 * (L / E))) void k_fft_rows_t(const C *__restrict__ data,
                                                          C *__restrict__ X,
                                                          const C *__restrict__ tw,
                                                          const double2 *__restrict__ winv,
                                                          int n1, long Mh, long M, int S,
                                                          SoiBatch b)
{
  constexpr int T = 8, TS = L / E;
  data = batch_ptr(data, b.tb);
  X = batch_ptr(X, b.x);
  using LY = QFast<L, T>;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  real_t<C> *ldsd = reinterpret_cast<real_t<C> *>(lds_raw);
  const int tid = threadIdx.x;
  const int q = tid % T, ts = tid / T;
  const int nrb = n1 / T;
  const int rb = blockIdx.x % nrb, s = blockIdx.x / nrb;
  if (s >= S) return;
  const int a = rb * T;
  const C *tile = data + (long)s * Mh + (long)a * L + q;

  C v[E];
  constexpr int R0 = RS::at(0);
#pragma unroll
  for (int c = 0; c < E / R0; ++c)
#pragma unroll
    for (int k = 0; k < R0; ++k) v[c * R0 + k] = tile[(unsigned)(ts + TS * c + k * (L / R0)) * 8u];

  run_stages<L, E, T, 1, TS, RS, LY, 0, C>(v, ldsd, q, ts, tw);

  constexpr int RL = RS::at(RS::n - 1);
  C *out = X + (long)s * M;
#pragma unroll
  for (int c = 0; c < E / RL; ++c)
#pragma unroll
    for (int k = 0; k < RL; ++k) {
      const long m = (long)(a + q) + (long)n1 * out_index<L, TS, RS>(ts, c, k);
      if (m < M) out[m] = demod<C>(winv[m], v[c * RL + k]);  // W_inv applied in FP64
    }
}

// ---------------------------------------------------------------------------------------------
// Pass 2: T adjacent rows of length L = n2, demodulated, stored transposed with 128 B runs
// ---------------------------------------------------------------------------------------------
template <int L, int E, int T, int NQP, int BPC, typename RS, typename C>
__global__ __launch_bounds__(T * (L / E), fft_minw(BPC, T * (L / E))) void k_fft_rows(const C *__restrict__ data,
                                                          C *__restrict__ X,
                                                          const C *__restrict__ tw,
                                                          const double2 *__restrict__ winv,
                                                          int n1, long Mh, long M, int S,
                                                          SoiBatch b)
{
  constexpr int TS = L / E;
  data = batch_ptr(data, b.at);
  X = batch_ptr(X, b.x);
  extern __shared__ __align__(16) unsigned char lds_raw[];
  real_t<C> *ldsd = reinterpret_cast<real_t<C> *>(lds_raw);
  const int tid = threadIdx.x;
  // loads: consecutive threads walk a row (coalesced); the sequence index is the slow one
  const int ts = tid % TS, q = tid / TS;
  const int nrb = n1 / T;
  const int bid = xcd_remap(blockIdx.x, gridDim.x);
  const int s = bid % S;  // the S segments of one row block share W_inv entries
  const int rb = bid / S;
  if (rb >= nrb) return;
  const int a = rb * T;
  const C *row = data + (long)s * Mh + (long)(a + q) * L;

  C v[E];
  constexpr int R0 = RS::at(0);
#pragma unroll
  for (int c = 0; c < E / R0; ++c)
#pragma unroll
    for (int k = 0; k < R0; ++k) v[c * R0 + k] = row[ts + TS * c + k * (L / R0)];

  using LY = SeqMajor<L, T / NQP>;
  run_stages<L, E, T, NQP, TS, RS, LY, 0, C>(v, ldsd, q, ts, tw);

  // transpose through LDS so that stores run over the T rows (m = a + q + n1*m2), in NQP groups
  // of T/NQP rows and real/imaginary halves
  constexpr int RL = RS::at(RS::n - 1);
  constexpr int TG = T / NQP;
  constexpr int NT = T * TS;
  constexpr int EG = E / NQP;  // values each thread stores per group
  C *out = X + (long)s * M;
  const int qg = q / TG, ql = q % TG;
#pragma unroll
  for (int g = 0; g < NQP; ++g) {
    C w[EG];
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      if (qg == g) {
#pragma unroll
        for (int c = 0; c < E / RL; ++c)
#pragma unroll
          for (int k = 0; k < RL; ++k)
            ldsd[LY::at(out_index<L, TS, RS>(ts, c, k), ql)] = half ? v[c * RL + k].y : v[c * RL + k].x;
      }
      lds_barrier();
#pragma unroll
      for (int c = 0; c < EG; ++c) {
        const int e = tid + c * NT;  // e = m2*TG + qq
        const real_t<C> val = ldsd[LY::at(e / TG, e % TG)];
        if (half) w[c].y = val; else w[c].x = val;
      }
      lds_barrier();
    }
#pragma unroll
    for (int c = 0; c < EG; ++c) {
      const int e = tid + c * NT;
      const long m = (long)(a + g * TG + e % TG) + (long)n1 * (e / TG);
      if (m < M) out[m] = demod<C>(winv[m], w[c]);  // W_inv applied in FP64
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Pass 2 with the sequence index fastest (as in the column pass): lanes q = tid % T of a wave
// hold T adjacent rows, so after the last stage each store instruction writes runs of T
// consecutive m (128 bytes at T = 8 complex128) straight from registers: no final LDS
// transpose.  Loads read T runs of T*16 bytes per instruction instead of one contiguous row run.
// ---------------------------------------------------------------------------------------------
template <int L, int E, int T, int NQP, int BPC, typename RS, typename C>
__global__ __launch_bounds__(T * (L / E), fft_minw(BPC, T * (L / E))) void k_fft_rows_q(const C *__restrict__ data,
                                                          C *__restrict__ X,
                                                          const C *__restrict__ tw,
                                                          const double2 *__restrict__ winv,
                                                          int n1, long Mh, long M, int S,
                                                          SoiBatch b)
{
  constexpr int TS = L / E;
  data = batch_ptr(data, b.at);
  X = batch_ptr(X, b.x);
  extern __shared__ __align__(16) unsigned char lds_raw[];
  real_t<C> *ldsd = reinterpret_cast<real_t<C> *>(lds_raw);
  const int tid = threadIdx.x;
  const int q = tid % T, ts = tid / T;
  const int nrb = n1 / T;
  const int bid = xcd_remap(blockIdx.x, gridDim.x);
  const int s = bid % S;  // the S segments of one row block share W_inv entries
  const int rb = bid / S;
  if (rb >= nrb) return;
  const int a = rb * T;
  const C *row = data + (long)s * Mh + (long)(a + q) * L;

  C v[E];
  constexpr int R0 = RS::at(0);
#pragma unroll
  for (int c = 0; c < E / R0; ++c)
#pragma unroll
    for (int k = 0; k < R0; ++k) v[c * R0 + k] = row[ts + TS * c + k * (L / R0)];

  run_stages<L, E, T, NQP, TS, RS, QFast<L, T / NQP>, 0, C>(v, ldsd, q, ts, tw);

  constexpr int RL = RS::at(RS::n - 1);
  C *out = X + (long)s * M;
#pragma unroll
  for (int c = 0; c < E / RL; ++c)
#pragma unroll
    for (int k = 0; k < RL; ++k) {
      const long m = (long)(a + q) + (long)n1 * out_index<L, TS, RS>(ts, c, k);
      if (m < M) out[m] = demod<C>(winv[m], v[c * RL + k]);  // W_inv applied in FP64
    }
}

// ---------------------------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------------------------
template <int L, int E, int T, int NQP, int BPC, typename RS, typename LY, typename C>
static int launch_cols(C *data, const C *tw, const C *thi, const C *tlo,
                       int n2, int lb1, long Mh, int S, C *tout, SoiBatch b, hipStream_t st)
{
  if (tout) {
    auto kt = k_fft_cols<L, E, T, NQP, BPC, RS, LY, C, true>;
    constexpr size_t lds = exch_lds_bytes<LY, C>();
    static bool once_t = false;
    if (!once_t) {
      (void)hipFuncSetAttribute((const void *)kt, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      once_t = true;
    }
    hipLaunchKernelGGL(kt, dim3((unsigned)((n2 / T) * S), b.n > 0 ? b.n : 1), dim3(T * (L / E)),
                       lds, st, data, tw, thi, tlo, n2, lb1, Mh, S, tout, b);
    return (int)hipGetLastError();
  }
  auto kern = k_fft_cols<L, E, T, NQP, BPC, RS, LY, C, false>;
  constexpr size_t lds = exch_lds_bytes<LY, C>();
  static_assert(lds <= 160 * 1024, "LDS");
  static bool once = false;
  if (!once) {
    (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    once = true;
  }
  hipLaunchKernelGGL(kern, dim3((unsigned)((n2 / T) * S), b.n > 0 ? b.n : 1), dim3(T * (L / E)),
                     lds, st, data, tw, thi, tlo, n2, lb1, Mh, S, (C *)nullptr, b);
  return (int)hipGetLastError();
}

template <int L, int E, int T, int NQP, int BPC, typename RS, typename C>
static int launch_rows(const C *data, C *X, const C *tw, const double2 *winv,
                       int n1, long Mh, long M, int S, SoiBatch b, hipStream_t st)
{
  auto kern = k_fft_rows<L, E, T, NQP, BPC, RS, C>;
  constexpr size_t lds = exch_lds_bytes<SeqMajor<L, T / NQP>, C>();
  static_assert(lds <= 160 * 1024, "LDS");
  static bool once = false;
  if (!once) {
    (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    once = true;
  }
  hipLaunchKernelGGL(kern, dim3((unsigned)((n1 / T) * S), b.n > 0 ? b.n : 1), dim3(T * (L / E)),
                     lds, st, data, X, tw, winv, n1, Mh, M, S, b);
  return (int)hipGetLastError();
}

template <int L, int E, int T, int NQP, int BPC, typename RS, typename C>
static int launch_rows_q(const C *data, C *X, const C *tw, const double2 *winv,
                         int n1, long Mh, long M, int S, SoiBatch b, hipStream_t st)
{
  auto kern = k_fft_rows_q<L, E, T, NQP, BPC, RS, C>;
  constexpr size_t lds = exch_lds_bytes<QFast<L, T / NQP>, C>();
  static_assert(lds <= 160 * 1024, "LDS");
  static bool once = false;
  if (!once) {
    (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    once = true;
  }
  hipLaunchKernelGGL(kern, dim3((unsigned)((n1 / T) * S), b.n > 0 ? b.n : 1), dim3(T * (L / E)),
                     lds, st, data, X, tw, winv, n1, Mh, M, S, b);
  return (int)hipGetLastError();
}

template <int L, int E, int BPC, typename RS, typename C>
static int launch_rows_t(const C *data, C *X, const C *tw, const double2 *winv, int n1, long Mh,
                         long M, int S, SoiBatch b, hipStream_t st)
{
  auto kern = k_fft_rows_t<L, E, BPC, RS, C>;
  constexpr size_t lds = exch_lds_bytes<QFast<L, 8>, C>();
  static_assert(lds <= 160 * 1024, "LDS");
  static bool once = false;
  if (!once) {
    (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    once = true;
  }
  hipLaunchKernelGGL(kern, dim3((unsigned)((n1 / 8) * S), b.n > 0 ? b.n : 1), dim3(8 * (L / E)),
                     lds, st, data, X, tw, winv, n1, Mh, M, S, b);
  return (int)hipGetLastError();
}

// Register-FFT lengths (E values per thread, radices in stage order):
//   256 = 16.16, 512 = 16.16.2, 1024 = 16.16.4, 2048 = 16.16.8            (E = 16)
//   320 = 10.4.4.2, 640 = 10.4.4.4, 1280 = 10.4.4.4.2, 2560 = 10.4.4.4.4   (E = 20)
// Columns take any of them (2560 exchanges in two sequence groups at 8-wide tiles); 2560-point
// rows run 4-wide tiles.
extern "C" int soik_fft_reg_supported(int L, int rows)
{
  switch (L) {
    case 256: case 320: case 512: case 640: case 1024: case 1280: case 2048:
      return 1;
    case 2560:
      return 1;
    case 20: case 30: case 40: case 50: case 60: case 80: case 100: case 160: case 200:
      return !rows;  // short columns only
    default:
      return 0;
  }
}

// Column tile width of the short column lengths (0: a regular length, tile 8 or 4)
extern "C" int soik_fft_cols_short_tile(int L)
{
  switch (L) {
    case 20: case 30: return 64;
    case 40: case 60: return 32;
    case 50: case 80: case 100: return 16;
    case 160: case 200: return 8;
    default: return 0;
  }
}

// Tile width T: 8 (128-byte runs, one 1024-thread block per CU for the long lengths) or 4
// (64-byte runs, two blocks per CU).
// Tile width T: 8 (128-byte runs) or 4 (64-byte runs).  variant selects a tuning alternative
// for some lengths: 1 = one padded LDS run per sequence (2048, 2560 at T = 8); 2 = the whole
// 256-VGPR budget for one 4-wide block per CU (2560 at T = 4); 3 = radices 4.4.4.4.10 (2560 at
// T = 8, complex64 only); 4 = one 4-wide block per CU exchanging whole complex128 values
// (QFull, 2560 at T = 4).
template <typename C>
static int cols_reg(C *data, const C *tw1, const C *thi, const C *tlo, int n1, int n2, int lb1,
                    long Mh, int S, int T, int variant, C *tout, SoiBatch b, hipStream_t st)
{
  if (n2 % T) return (int)hipErrorInvalidValue;
#define COLS(VV, LV, EV, TV, NQ, BPC, LY, ...) \
  if (n1 == LV && T == TV && variant == VV) \
    return launch_cols<LV, EV, TV, NQ, BPC, Radices<__VA_ARGS__>, LY<LV, TV / NQ>, C>(data, tw1, thi, tlo, n2, lb1, Mh, S, tout, b, st);
  COLS(0, 256, 16, 8, 1, 1, QFast, 16, 16) COLS(0, 320, 20, 8, 1, 1, QFast, 10, 4, 4, 2)
  COLS(0, 512, 16, 8, 1, 1, QFast, 16, 16, 2) COLS(0, 640, 20, 8, 1, 1, QFast, 10, 4, 4, 4)
  COLS(0, 1024, 16, 8, 1, 1, QFast, 16, 16, 4) COLS(0, 1280, 20, 8, 1, 1, QFast, 10, 4, 4, 4, 2)
  COLS(0, 2048, 16, 8, 1, 1, QFast, 16, 16, 8) COLS(0, 2560, 20, 8, 2, 1, QFast8, 10, 4, 4, 4, 4)
  COLS(0, 1024, 16, 4, 1, 2, QFast, 16, 16, 4) COLS(0, 1280, 20, 4, 1, 2, QFast, 10, 4, 4, 4, 2)
  COLS(0, 2048, 16, 4, 1, 2, QFast, 16, 16, 8) COLS(0, 2560, 20, 4, 2, 2, QFast, 10, 4, 4, 4, 4)
  COLS(1, 2048, 16, 8, 1, 1, SeqMajor, 16, 16, 8) COLS(1, 2560, 20, 8, 2, 1, SeqMajor, 10, 4, 4, 4, 4)
  COLS(2, 2560, 20, 4, 1, 1, QFast, 10, 4, 4, 4, 4)
  if constexpr (sizeof(C) == 16) {  // complex128: whole-value exchanges in the full 160 KiB
    COLS(4, 2560, 20, 4, 1, 1, QFull, 10, 4, 4, 4, 4)
  }
  if constexpr (sizeof(C) == 8) {  // complex64: the radix-10 stage last needs fewer registers
    COLS(3, 2560, 20, 8, 2, 1, QFast, 4, 4, 4, 4, 10)
  }
  // short columns (M_hat = n1 x 256 with n1 of 2^a 3^b 5^c form): wide tiles, one block of
  // 64 - 80 threads over 8 - 64 columns
  COLS(0, 20, 20, 64, 1, 1, QFast, 10, 2) COLS(0, 30, 30, 64, 1, 1, QFast, 10, 3)
  COLS(0, 40, 20, 32, 1, 1, QFast, 10, 4) COLS(0, 50, 10, 16, 1, 1, QFast, 10, 5)
  COLS(0, 60, 30, 32, 1, 1, QFast, 10, 3, 2) COLS(0, 80, 20, 16, 1, 1, QFast, 10, 4, 2)
  COLS(0, 100, 20, 16, 1, 1, QFast, 10, 10) COLS(0, 160, 20, 8, 1, 1, QFast, 10, 4, 4)
  COLS(0, 200, 20, 8, 1, 1, QFast, 10, 10, 2)
#undef COLS
  return (int)hipErrorInvalidValue;
}

template <typename C>
static int rows_reg(const C *data, C *X, const C *tw2, const double2 *winv, int n1, int n2,
                    long Mh, long M, int S, int T, SoiBatch b, hipStream_t st)
{
  if (n1 % T) return (int)hipErrorInvalidValue;
#define ROWS(LV, EV, TV, NQ, BPC, ...) \
  if (n2 == LV && T == TV) return launch_rows<LV, EV, TV, NQ, BPC, Radices<__VA_ARGS__>, C>(data, X, tw2, winv, n1, Mh, M, S, b, st);
  // 8-row tiles of 1024 and 2048 points: the sequence-index-fastest kernel, which stores from
  // registers (no final transpose; 2048 complex128: 1.37 -> 1.29 ms).  SOI_FFT_ROWSQ=0 selects
  // the row-walking kernel.
  {
    static const int rq = [] {
      const char *e = getenv("SOI_FFT_ROWSQ");
      return e ? atoi(e) : 1;
    }();
    if (rq && T == 8 && n2 == 2048)
      return launch_rows_q<2048, 16, 8, 1, 1, Radices<16, 16, 8>, C>(data, X, tw2, winv, n1, Mh, M, S, b, st);
    if (rq && T == 8 && n2 == 1024)
      return launch_rows_q<1024, 16, 8, 1, 1, Radices<16, 16, 4>, C>(data, X, tw2, winv, n1, Mh, M, S, b, st);
  }
  ROWS(256, 16, 8, 1, 1, 16, 16) ROWS(320, 20, 8, 1, 1, 10, 4, 4, 2) ROWS(512, 16, 8, 1, 1, 16, 16, 2)
  ROWS(640, 20, 8, 1, 1, 10, 4, 4, 4) ROWS(1024, 16, 8, 1, 1, 16, 16, 4)
  ROWS(1280, 20, 8, 1, 1, 10, 4, 4, 4, 2) ROWS(2048, 16, 8, 1, 1, 16, 16, 8)
  ROWS(2560, 20, 4, 1, 1, 10, 4, 4, 4, 4)
  ROWS(1024, 16, 4, 1, 2, 16, 16, 4) ROWS(1280, 20, 4, 1, 2, 10, 4, 4, 4, 2)
  ROWS(2048, 16, 4, 1, 2, 16, 16, 8)
  ROWS(256, 16, 10, 1, 1, 16, 16)  // rows over short columns whose count is not a multiple of 8
#undef ROWS
  return (int)hipErrorInvalidValue;
}

template <typename C>
static int rows_t(const C *data, C *X, const C *tw2, const double2 *winv, int n1, int n2, long Mh,
                  long M, int S, SoiBatch b, hipStream_t st)
{
  if (n1 % 8) return (int)hipErrorInvalidValue;
#define ROWST(LV, EV, BPC, ...) \
  if (n2 == LV) return launch_rows_t<LV, EV, BPC, Radices<__VA_ARGS__>, C>(data, X, tw2, winv, n1, Mh, M, S, b, st);
  ROWST(256, 16, 1, 16, 16) ROWST(320, 20, 1, 10, 4, 4, 2) ROWST(512, 16, 1, 16, 16, 2)
  ROWST(640, 20, 1, 10, 4, 4, 4) ROWST(1024, 16, 1, 16, 16, 4)
  ROWST(1280, 20, 1, 10, 4, 4, 4, 2) ROWST(2048, 16, 1, 16, 16, 8)
  if constexpr (sizeof(C) == 8) { ROWST(2560, 20, 1, 10, 4, 4, 4, 4) }
#undef ROWST
  return (int)hipErrorInvalidValue;
}

extern "C" int soik_fft_rows_tiled_supported(int n2, int f32)
{
  switch (n2) {
    case 256: case 320: case 512: case 640: case 1024: case 1280: case 2048: return 1;
    case 2560: return f32;
    default: return 0;
  }
}

extern "C" int soik_fft_cols_reg(double2 *data, const double2 *tw1, const double2 *thi,
                                 const double2 *tlo, int n1, int n2, int lb1, long Mh, int S,
                                 int T, int variant, double2 *tout, SoiBatch b, hipStream_t st)
{
  return cols_reg<double2>(data, tw1, thi, tlo, n1, n2, lb1, Mh, S, T, variant, tout, b, st);
}

extern "C" int soik_fft_rows_tiled(const double2 *data, double2 *X, const double2 *tw2,
                                   const double2 *winv, int n1, int n2, long Mh, long M, int S,
                                   SoiBatch b, hipStream_t st)
{
  return rows_t<double2>(data, X, tw2, winv, n1, n2, Mh, M, S, b, st);
}

extern "C" int soik_fft_rows_tiled_f32(const float2 *data, float2 *X, const float2 *tw2,
                                       const double2 *winv, int n1, int n2, long Mh, long M,
                                       int S, SoiBatch b, hipStream_t st)
{
  return rows_t<float2>(data, X, tw2, winv, n1, n2, Mh, M, S, b, st);
}

extern "C" int soik_fft_rows_reg(const double2 *data, double2 *X, const double2 *tw2,
                                 const double2 *winv, int n1, int n2, long Mh, long M, int S,
                                 int T, SoiBatch b, hipStream_t st)
{
  return rows_reg<double2>(data, X, tw2, winv, n1, n2, Mh, M, S, T, b, st);
}

// FP32 form: complex64 data and tables (rounded from their FP64 values); W_inv stays FP64.
extern "C" int soik_fft_cols_reg_f32(float2 *data, const float2 *tw1, const float2 *thi,
                                     const float2 *tlo, int n1, int n2, int lb1, long Mh, int S,
                                     int T, int variant, float2 *tout, SoiBatch b, hipStream_t st)
{
  return cols_reg<float2>(data, tw1, thi, tlo, n1, n2, lb1, Mh, S, T, variant, tout, b, st);
}

extern "C" int soik_fft_rows_reg_f32(const float2 *data, float2 *X, const float2 *tw2,
                                     const double2 *winv, int n1, int n2, long Mh, long M, int S,
                                     int T, SoiBatch b, hipStream_t st)
{
  return rows_reg<float2>(data, X, tw2, winv, n1, n2, Mh, M, S, T, b, st);
}
