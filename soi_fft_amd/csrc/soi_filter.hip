// soi_filter.hip -- Step 1 of the SOI FFT on MI355X: B-tap polyphase filter, subsampling by
// n_mu/d_mu and the S-point DFT across lanes, stored segment-major (the single-GPU all-to-all).
//
// Reference: parallel_filter_subsampling.cpp:52-527 (interior conv 134-287, S-point DFTs
// 297-311, transpose 316-395, tail rows 403-446, halo rows 496-524; the halo exchange 101-108
// becomes circular indexing of the input).
//
//   gamma[R][i] = sum_{b<B} w[theta][b][i] * alpha[((J*d + b) mod M)*S + i],  R = J*n + theta
//   alpha_tilde[s][R] = sum_{i<S} gamma[R][i] * e^{-2 pi i i s / S}
//
// Work tile: U*W consecutive row groups J (all S lanes, NT of the n_mu phases per block).
// Thread (i, theta, p, w) keeps BP taps of lane i / phase theta / tap slice p in registers for
// the whole kernel and accumulates U outputs of J-stream w from an LDS copy of the tile's input
// rows.  Real-tap form (CT = false): w = post[theta][i] * taps with real taps (exact windows,
// the e^{i pi t} factor splits into a per-(theta,i) phase and (-1)^b).  Complex-tap form
// (CT = true): the reference's complex window values, 4 FMAs per tap.
//
// Per tile: the fully unrolled FMA block over an LDS copy of the tile's contiguous input rows,
// the tap-slice reduction through LDS, the post-phase multiply, an LDS Stockham S-point DFT per
// row, and coalesced stores of n_mu*U*W consecutive R per segment.  The next tile's rows are
// loaded into registers while this tile computes, and barriers wait on LDS traffic only.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "soi_common.hpp"
#include "soi_kernels.h"

using namespace soi;

typedef __attribute__((address_space(3))) void lds_void_t;


template <int BP, bool CT> struct Taps;
template <int BP> struct Taps<BP, false> { double v[BP]; };
template <int BP> struct Taps<BP, true> { double re[BP], im[BP]; };

// Fully unrolled accumulation: input row o of the thread's column feeds accumulator u with tap
// b = o - u*d whenever 0 <= b < BP (constant indices: taps and accumulators stay in registers).
template <int O, int UU, int U, int DMU, int BP, bool CT>
__device__ __forceinline__ void conv_taps(const double2 xv, const Taps<BP, CT> &h, double2 (&acc)[U])
{
  if constexpr (UU < U) {
    constexpr int b = O - UU * DMU;
    if constexpr (b >= 0 && b < BP) {
      if constexpr (CT) {
        acc[UU].x = fma(h.re[b], xv.x, acc[UU].x);
        acc[UU].x = fma(-h.im[b], xv.y, acc[UU].x);
        acc[UU].y = fma(h.re[b], xv.y, acc[UU].y);
        acc[UU].y = fma(h.im[b], xv.x, acc[UU].y);
      } else {
        acc[UU].x = fma(h.v[b], xv.x, acc[UU].x);
        acc[UU].y = fma(h.v[b], xv.y, acc[UU].y);
      }
    }
    conv_taps<O, UU + 1, U, DMU, BP, CT>(xv, h, acc);
  }
}

template <int O, int OEND, int S, int U, int DMU, int BP, bool CT>
__device__ __forceinline__ void conv_rows(const double2 *__restrict__ col, const Taps<BP, CT> &h,
                                          double2 (&acc)[U])
{
  if constexpr (O < OEND) {
    const double2 xv = col[O * S];
    conv_taps<O, 0, U, DMU, BP, CT>(xv, h, acc);
    conv_rows<O + 1, OEND, S, U, DMU, BP, CT>(col, h, acc);
  }
}

template <int S, int NMU, int NT, int DMU, int BP, int U, int NP, int W, bool CT>
struct FilterCfg {
  static constexpr int T = S * NT * NP * W;          // threads
  static constexpr int NWAVES = T / 64;
  static constexpr int TJ = U * W;                   // row groups per tile
  static constexpr int ROWS = TJ * DMU + NP * BP - DMU;  // input rows per tile
  static constexpr int XN = (ROWS * S + 63) / 64 * 64;   // staged complex values (whole waves)
  static constexpr int NDMA = XN / 64;               // wave-wide 16-byte loads per tile
  static constexpr int GR = TJ * NT;                 // output rows per tile (this block's phases)
  static constexpr int GS = S + 1;                   // padded G row stride
  static constexpr int GN = GR * GS;
  static constexpr int BUFN = (XN > GN ? XN : GN);   // one tile buffer (G reuses it)
  static constexpr int NG = NMU / NT;                // phase groups
  // two tile buffers (the next tile's rows stream in while this one computes) when they fit
  static constexpr bool DB = 2 * BUFN * 16 + S * 16 <= 160 * 1024;
  static constexpr int LDS_BYTES = ((DB ? 2 : 1) * BUFN + S) * 16;
  // blocks resident per CU (LDS and registers), and the register budget that follows
  static constexpr int BPC = (LDS_BYTES <= 80 * 1024 && T <= 320) ? 2 : 1;
  static constexpr int MINW = (BPC * NWAVES + 3) / 4;
  static constexpr int NSC = (GR * S + T - 1) / T;   // stores per thread per tile (at most)
  static constexpr int NSC_MIN = (GR * S) / T;       // (at least)
};

constexpr int filter_threads(int S, int NT, int NP, int W) { return S * NT * NP * W; }

// Barrier for LDS hand-off only: waits for this wave's LDS operations, not for its global
// stores, which stay in flight across it.
__device__ __forceinline__ void lds_barrier()
{
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// s_waitcnt vmcnt(n) (expcnt and lgkmcnt left alone), through the builtin so the compiler's
// own wait tracking sees it
template <int n> __device__ __forceinline__ void wait_vmcnt()
{
  static_assert(n >= 0 && n < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((n & 15) | (7 << 4) | (15 << 8) | ((n >> 4) << 14));
}

// In-place S-point DFT of each of the GR rows of G (row r at G + r*GS, unit stride), Stockham
// passes of radix 8 then 4 / 2 with compile-time shapes; tw[m] = e^{-2 pi i m/S} in LDS.
template <int S> constexpr int sdft_radix(int pass)
{
  int rem = S, p = 0, r = 0;
  while (rem > 1) {
    r = (rem % 8 == 0 && rem != 16) ? 8 : (rem % 4 == 0 ? 4 : 2);
    if (p == pass) return r;
    rem /= r;
    ++p;
  }
  return 0;
}

template <int S, int GR, int GS, int T, int P, int NS>
__device__ __forceinline__ void sdft_passes(double2 *G, const double2 *tw, int tid)
{
  constexpr int R = sdft_radix<S>(P);
  if constexpr (R != 0) {
    constexpr int NBF = S / R;                 // butterflies per row
    constexpr int TOT = GR * NBF;
    constexpr int NB = (TOT + T - 1) / T;
    double2 v[NB][R];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      const int g = tid + c * T;
      if (TOT % T == 0 || g < TOT) {
        const int q = g / NBF, j = g % NBF;
        const double2 *src = G + q * GS + j;
#pragma unroll
        for (int k = 0; k < R; ++k) v[c][k] = src[k * NBF];
      }
    }
    lds_barrier();
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      const int g = tid + c * T;
      if (TOT % T == 0 || g < TOT) {
        const int q = g / NBF, j = g % NBF;
        const int jm = j % NS;
        if constexpr (NS > 1) {
          constexpr int step = S / (NS * R);
#pragma unroll
          for (int k = 1; k < R; ++k) v[c][k] = cmul(v[c][k], tw[k * jm * step]);
        }
        dft_r<R>(v[c]);
        double2 *dst = G + q * GS + (j - jm) * R + jm;
#pragma unroll
        for (int k = 0; k < R; ++k) dst[k * NS] = v[c][k];
      }
    }
    lds_barrier();
    sdft_passes<S, GR, GS, T, P + 1, NS * R>(G, tw, tid);
  }
}

// Stage the XN input values of a tile (contiguous from element g0, wrapping once at N) into
// dst with direct-to-LDS loads; wave w moves 64-value chunks w, w + NWAVES, ...
template <typename C>
__device__ __forceinline__ void dma_rows(double2 *dst, const double2 *__restrict__ x, long g0,
                                         long N, int wave, int lane)
{
#pragma unroll
  for (int c = wave; c < C::NDMA; c += C::NWAVES) {
    long gi = g0 + c * 64 + lane;
    if (gi >= N) gi -= N;
    __builtin_amdgcn_global_load_lds((const void *)(x + gi), (lds_void_t *)(dst + c * 64), 16,
                                     0, 0);
  }
}

// Per-thread state that stays in registers for the whole kernel.
template <int BP, bool CT> struct FilterThread {
  Taps<BP, CT> h;
  double2 pf;        // post phase
  int i, tl, p, w, g;
  int tid, wave, lane;
};

// One tile: prefetch the next tile's rows into nxt (when there is a next tile), FMA block from
// cur, tap-slice reduction + post phase + S-point DFT in cur, stores.  Barriers wait on LDS
// traffic only: the prefetch and the stores stay in flight.
template <typename C, int S, int NMU, int NT, int DMU, int BP, int U, int NP, bool CT>
__device__ __forceinline__ void filter_tile(const FilterThread<BP, CT> &th, double2 *cur,
                                            double2 *nxt, const double2 *tw,
                                            const double2 *__restrict__ x,
                                            double2 *__restrict__ at, long tile, bool more,
                                            long N, long Mh, long nJ)
{
  const long J0 = tile * C::TJ;
  if constexpr (C::DB) {
    if (more) dma_rows<C>(nxt, x, (tile + 1) * C::TJ * DMU * S, N, th.wave, th.lane);
  } else {
    dma_rows<C>(cur, x, tile * C::TJ * DMU * S, N, th.wave, th.lane);
    wait_vmcnt<0>();
    lds_barrier();
  }

  // ---- FMA block
  double2 acc[U];
#pragma unroll
  for (int u = 0; u < U; ++u) acc[u] = double2{0.0, 0.0};
  conv_rows<0, (U - 1) * DMU + BP, S, U, DMU, BP, CT>(cur + (th.w * U * DMU + th.p * BP) * S + th.i,
                                                      th.h, acc);
  lds_barrier();

  // ---- tap-slice reduction (fixed order p = NP-1 .. 0) and post phase, into G = cur
  double2 *G = cur;
  const int gbase = (th.w * U) * NT + th.tl;
#pragma unroll
  for (int q = NP - 1; q >= 0; --q) {
    if (th.p == q) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        double2 a = acc[u];
        double2 *gp = G + (gbase + u * NT) * C::GS + th.i;
        if (q != NP - 1) a = cadd(a, *gp);
        if (q == 0 && !CT) a = cmul(a, th.pf);
        *gp = a;
      }
    }
    lds_barrier();
  }

  // ---- S-point DFT of each row
  sdft_passes<S, C::GR, C::GS, C::T, 0, 1>(G, tw, th.tid);

  // ---- store segment-major: R = J*n_mu + theta (r fastest: runs of consecutive R)
#pragma unroll
  for (int c = 0; c < C::NSC; ++c) {
    const int e = th.tid + c * C::T;
    if ((C::GR * S) % C::T == 0 || e < C::GR * S) {
      const int s = e / C::GR, r = e % C::GR;
      const long J = J0 + r / NT;
      if (J < nJ) {
        const long R = J * NMU + th.g * NT + r % NT;
        at[(long)s * Mh + R] = G[r * C::GS + s];
      }
    }
  }
  if constexpr (C::DB) {
    // this wave's prefetch loads are complete once at most its stores are outstanding (the
    // stores were issued after them; a tile with a successor stores every row)
    asm volatile("" ::: "memory");
    if (more) wait_vmcnt<C::NSC_MIN>();
  }
  lds_barrier();
}

template <int S, int NMU, int NT, int DMU, int BP, int U, int NP, int W, bool CT>
__global__ __launch_bounds__(filter_threads(S, NT, NP, W),
                             (FilterCfg<S, NMU, NT, DMU, BP, U, NP, W, CT>::MINW)) void k_filter(
    const double2 *__restrict__ x, double2 *__restrict__ at, const double *__restrict__ taps,
    const double2 *__restrict__ post, const double2 *__restrict__ twS, long M, long Mh, long nJ,
    int tiles_per_block)
{
  typedef FilterCfg<S, NMU, NT, DMU, BP, U, NP, W, CT> C;
  __shared__ double2 bufA[C::BUFN];
  __shared__ double2 bufB[C::DB ? C::BUFN : 1];
  __shared__ double2 tw[S];
  FilterThread<BP, CT> th;
  th.tid = threadIdx.x;
  th.wave = th.tid >> 6;
  th.lane = th.tid & 63;
  th.i = th.tid % S;
  int rest = th.tid / S;
  th.tl = rest % NT;
  rest /= NT;
  th.p = rest % NP;
  th.w = rest / NP;

  // phase group: the slow index of the block id so blocks sharing input rows are 8-apart
  // (same XCD under round-robin dispatch) and walk the same tiles at the same time
  const int nblk_g = gridDim.x / C::NG;
  th.g = blockIdx.x / nblk_g;
  const int bg = blockIdx.x - th.g * nblk_g;
  const int theta = th.g * NT + th.tl;
  {
    const long base = ((long)theta * S + th.i) * (NP * BP) + th.p * BP;
    if constexpr (CT) {
      const double2 *tp = reinterpret_cast<const double2 *>(taps) + base;
#pragma unroll
      for (int b = 0; b < BP; ++b) {
        double2 v = tp[b];
        th.h.re[b] = v.x;
        th.h.im[b] = v.y;
      }
    } else {
      const double *tp = taps + base;
#pragma unroll
      for (int b = 0; b < BP; ++b) th.h.v[b] = tp[b];
    }
  }
  th.pf = post[theta * S + th.i];
  for (int e = th.tid; e < S; e += C::T) tw[e] = twS[e];

  const long ntiles = (nJ + C::TJ - 1) / C::TJ;
  const long t0 = (long)bg * tiles_per_block;
  const long t1 = min(ntiles, t0 + tiles_per_block);
  const long N = M * S;

  if constexpr (C::DB) {
    if (t0 < t1) dma_rows<C>(bufA, x, t0 * C::TJ * DMU * S, N, th.wave, th.lane);
    wait_vmcnt<0>();
    lds_barrier();
    for (long tile = t0; tile < t1; ++tile) {
      const bool odd = (tile - t0) & 1;
      filter_tile<C, S, NMU, NT, DMU, BP, U, NP, CT>(th, odd ? bufB : bufA, odd ? bufA : bufB,
                                                      tw, x, at, tile, tile + 1 < t1, N, Mh, nJ);
    }
  } else {
    lds_barrier();
    for (long tile = t0; tile < t1; ++tile)
      filter_tile<C, S, NMU, NT, DMU, BP, U, NP, CT>(th, bufA, bufA, tw, x, at, tile, false, N,
                                                      Mh, nJ);
  }
}

// ---------------------------------------------------------------------------------------------
// Dispatch table
// ---------------------------------------------------------------------------------------------
template <int S, int NMU, int NT, int DMU, int BP, int U, int NP, int W, bool CT>
static int launch(const SoiFilterArgs *a, hipStream_t st)
{
  typedef FilterCfg<S, NMU, NT, DMU, BP, U, NP, W, CT> C;
  static_assert(C::LDS_BYTES <= 160 * 1024, "filter tile exceeds LDS");
  if ((long)C::XN > a->M * S) return (int)hipErrorInvalidValue;  // tile window wraps at most once
  auto kern = k_filter<S, NMU, NT, DMU, BP, U, NP, W, CT>;
  const long ntiles = (a->nJ + C::TJ - 1) / C::TJ;
  // blocks per phase group: enough to fill the CUs, tiles split into contiguous ranges
  long nb = 256L * C::BPC / C::NG;
  if (nb < 1) nb = 1;
  if (nb > ntiles) nb = ntiles;
  const int tpb = (int)((ntiles + nb - 1) / nb);
  nb = (ntiles + tpb - 1) / tpb;
  hipLaunchKernelGGL(kern, dim3((unsigned)(nb * C::NG)), dim3(C::T), 0, st,
                     (const double2 *)a->x, (double2 *)a->at, a->taps, a->post,
                     (const double2 *)a->twS, a->M, a->Mh, a->nJ, tpb);
  return (int)hipGetLastError();
}

// Shape for (S, n_mu, d_mu, B, complex taps).  Returns 0 and fills NP/BP/NT/U/W, or -1.
extern "C" int soik_filter_shape(int S, int n_mu, int d_mu, int B, int ct, int *NP, int *BP,
                                 int *NT, int *U, int *W)
{
  if (!ct) {
    if (n_mu != 5 || d_mu != 4) return -1;
    *NP = 2;
    *U = 8;
    *NT = 5;
    int half = (B + 1) / 2;
    *BP = half <= 32 ? 32 : half <= 36 ? 36 : half <= 40 ? 40 : 0;
    if (!*BP || B < 60) return -1;
    switch (S) {  /* 640 threads: S * 5 phases * 2 tap slices * W streams */
      case 4: *W = 16; break;
      case 8: *W = 8; break;
      case 16: *W = 4; break;
      case 32: *W = 2; break;
      case 64: *W = 1; break;
      default: return -1;
    }
    return 0;
  }
  *NP = 4;
  int q = (B + 3) / 4;
  if (n_mu == 5 && d_mu == 4) {
    *U = 8;
    *BP = q <= 8 ? 8 : q <= 12 ? 12 : q <= 16 ? 16 : q <= 20 ? 20 : 0;
    if (!*BP) return -1;
    switch (S) {  /* S * NT * 4 tap slices * W threads */
      case 4: *NT = 5; *W = 8; break;
      case 8: *NT = 5; *W = 4; break;
      case 16: *NT = 5; *W = 2; break;
      case 32: *NT = 5; *W = 1; break;
      case 64: *NT = 1; *W = 2; break;
      default: return -1;
    }
    return 0;
  }
  if (n_mu == 8 && d_mu == 7) {
    *U = 8;
    *BP = q <= 16 ? 16 : q <= 20 ? 20 : 0;
    if (!*BP) return -1;
    switch (S) {
      case 4: *NT = 8; *W = 2; break;
      case 8: *NT = 8; *W = 1; break;
      case 16: *NT = 8; *W = 1; break;
      case 32: *NT = 4; *W = 1; break;
      case 64: *NT = 2; *W = 1; break;
      default: return -1;
    }
    return 0;
  }
  return -1;
}

extern "C" size_t soik_filter_lds_bytes(const SoiFilterArgs *a)
{
  const int TJ = a->U * a->W;
  const int rows = TJ * a->d_mu + a->NP * a->BP - a->d_mu;
  const long xn = ((long)rows * a->S + 63) / 64 * 64;
  const long gn = (long)TJ * a->NT * (a->S + 1);
  const long buf = xn > gn ? xn : gn;
  const int db = 2 * buf * 16 + a->S * 16 <= 160 * 1024;
  return (size_t)((db ? 2 : 1) * buf + a->S) * sizeof(double2);
}

#define RT_CASE(SV, BPV, WV)                                                        \
  if (a->S == SV && a->BP == BPV && a->W == WV) return launch<SV, 5, 5, 4, BPV, 8, 2, WV, false>(a, st);
#define CT54_CASE(SV, NTV, BPV, WV)                                                 \
  if (a->S == SV && a->BP == BPV) return launch<SV, 5, NTV, 4, BPV, 8, 4, WV, true>(a, st);
#define CT87_CASE(SV, NTV, BPV, WV)                                                 \
  if (a->S == SV && a->BP == BPV) return launch<SV, 8, NTV, 7, BPV, 8, 4, WV, true>(a, st);

extern "C" int soik_filter_sdft(const SoiFilterArgs *a, hipStream_t st)
{
  if (!a->ct) {
    RT_CASE(4, 32, 16) RT_CASE(4, 36, 16) RT_CASE(4, 40, 16)
    RT_CASE(8, 32, 8) RT_CASE(8, 36, 8) RT_CASE(8, 40, 8)
    RT_CASE(16, 32, 4) RT_CASE(16, 36, 4) RT_CASE(16, 40, 4)
    RT_CASE(32, 32, 2) RT_CASE(32, 36, 2) RT_CASE(32, 40, 2)
    RT_CASE(64, 32, 1) RT_CASE(64, 36, 1) RT_CASE(64, 40, 1)
  } else if (a->n_mu == 5) {
    CT54_CASE(4, 5, 8, 8) CT54_CASE(4, 5, 12, 8) CT54_CASE(4, 5, 16, 8) CT54_CASE(4, 5, 20, 8)
    CT54_CASE(8, 5, 8, 4) CT54_CASE(8, 5, 12, 4) CT54_CASE(8, 5, 16, 4) CT54_CASE(8, 5, 20, 4)
    CT54_CASE(16, 5, 8, 2) CT54_CASE(16, 5, 12, 2) CT54_CASE(16, 5, 16, 2) CT54_CASE(16, 5, 20, 2)
    CT54_CASE(32, 5, 8, 1) CT54_CASE(32, 5, 12, 1) CT54_CASE(32, 5, 16, 1) CT54_CASE(32, 5, 20, 1)
    CT54_CASE(64, 1, 8, 2) CT54_CASE(64, 1, 12, 2) CT54_CASE(64, 1, 16, 2) CT54_CASE(64, 1, 20, 2)
  } else {
    CT87_CASE(4, 8, 16, 2) CT87_CASE(4, 8, 20, 2)
    CT87_CASE(8, 8, 16, 1) CT87_CASE(8, 8, 20, 1)
    CT87_CASE(16, 8, 16, 1) CT87_CASE(16, 8, 20, 1)
    CT87_CASE(32, 4, 16, 1) CT87_CASE(32, 4, 20, 1)
    CT87_CASE(64, 2, 16, 1) CT87_CASE(64, 2, 20, 1)
  }
  return (int)hipErrorInvalidValue;
}
