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
// Per tile: direct-to-LDS loads (global_load_lds_dwordx4) of the contiguous input rows, the
// fully unrolled FMA block, the tap-slice reduction through LDS, the post-phase multiply, an LDS
// Stockham S-point DFT per row, and coalesced stores of n_mu*U*W consecutive R per segment.
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
  static constexpr int TJ = U * W;                   // row groups per tile
  static constexpr int ROWS = TJ * DMU + NP * BP - DMU;  // input rows per tile
  static constexpr int XN = (ROWS * S + 63) / 64 * 64;   // staged complex values (whole waves)
  static constexpr int GR = TJ * NT;                 // output rows per tile (this block's phases)
  static constexpr int GS = S + 1;                   // padded G row stride
  static constexpr int GN = GR * GS;
  static constexpr int LDSN = (XN > GN ? XN : GN);   // complex values; + S twiddles after
  static constexpr int NG = NMU / NT;                // phase groups
  // register budget: two blocks per CU for blocks of up to 6 waves, else one
  static constexpr int BLK_PER_CU = (T <= 384) ? 2 : 1;
  static constexpr int MINW = (BLK_PER_CU * (T / 64) + 3) / 4;
};

constexpr int filter_threads(int S, int NT, int NP, int W) { return S * NT * NP * W; }
constexpr int filter_minw(int T) { return (((T <= 384) ? 2 : 1) * (T / 64) + 3) / 4; }

template <int S, int NMU, int NT, int DMU, int BP, int U, int NP, int W, bool CT>
__global__ __launch_bounds__(filter_threads(S, NT, NP, W),
                             filter_minw(filter_threads(S, NT, NP, W))) void k_filter(
    const double2 *__restrict__ x, double2 *__restrict__ at, const double *__restrict__ taps,
    const double2 *__restrict__ post, const double2 *__restrict__ twS, long M, long Mh, long nJ,
    int tiles_per_block)
{
  typedef FilterCfg<S, NMU, NT, DMU, BP, U, NP, W, CT> C;
  extern __shared__ double2 lds[];
  double2 *X = lds;
  double2 *G = lds;  // reuses the input rows once the FMA block is done
  double2 *tw = lds + C::LDSN;
  const int tid = threadIdx.x;
  const int i = tid % S;
  int rest = tid / S;
  const int tl = rest % NT;
  rest /= NT;
  const int p = rest % NP;
  const int w = rest / NP;
  const int lane = tid & 63, wave = tid >> 6;
  constexpr int NWAVES = C::T / 64;

  // phase group: the slow index of the block id so blocks sharing input rows are 8-apart
  // (same XCD under round-robin dispatch) and walk the same tiles at the same time
  const int nblk_g = gridDim.x / C::NG;
  const int g = blockIdx.x / nblk_g;
  const int bg = blockIdx.x - g * nblk_g;
  const int theta = g * NT + tl;

  Taps<BP, CT> h;
  {
    const long base = ((long)theta * S + i) * (NP * BP) + p * BP;
    if constexpr (CT) {
      const double2 *tp = reinterpret_cast<const double2 *>(taps) + base;
#pragma unroll
      for (int b = 0; b < BP; ++b) {
        double2 v = tp[b];
        h.re[b] = v.x;
        h.im[b] = v.y;
      }
    } else {
      const double *tp = taps + base;
#pragma unroll
      for (int b = 0; b < BP; ++b) h.v[b] = tp[b];
    }
  }
  const double2 pf = post[theta * S + i];
  for (int e = tid; e < S; e += C::T) tw[e] = twS[e];

  const long ntiles = (nJ + C::TJ - 1) / C::TJ;
  const long t0 = (long)bg * tiles_per_block;
  const long t1 = min(ntiles, t0 + tiles_per_block);
  const long N = M * S;

  for (long tile = t0; tile < t1; ++tile) {
    const long J0 = tile * C::TJ;
    // ---- stage the tile's input rows: contiguous in HBM (rows of S lanes), circular at M
    {
      const long g0 = J0 * DMU * S;
#pragma unroll 4
      for (int c = wave; c < C::XN / 64; c += NWAVES) {
        long gi = g0 + c * 64 + lane;
        if (gi >= N) gi -= N;
        __builtin_amdgcn_global_load_lds((const void *)(x + gi), (lds_void_t *)(X + c * 64), 16,
                                         0, 0);
      }
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }
    __syncthreads();

    // ---- FMA block
    double2 acc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) acc[u] = double2{0.0, 0.0};
    conv_rows<0, (U - 1) * DMU + BP, S, U, DMU, BP, CT>(X + (w * U * DMU + p * BP) * S + i, h, acc);
    __syncthreads();

    // ---- tap-slice reduction (fixed order p = NP-1 .. 0) and post phase
    const int gbase = (w * U) * NT + tl;
    for (int q = NP - 1; q >= 0; --q) {
      if (p == q) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          double2 a = acc[u];
          double2 *gp = G + (gbase + u * NT) * C::GS + i;
          if (q != NP - 1) a = cadd(a, *gp);
          if (q == 0 && !CT) a = cmul(a, pf);
          *gp = a;
        }
      }
      if (q > 0) __syncthreads();
    }
    __syncthreads();

    // ---- S-point DFT of each row
    lds_fft<8>(G, S, C::GR, C::GS, 1, tw);

    // ---- store segment-major: R = J*n_mu + theta
#pragma unroll 2
    for (int e = tid; e < C::GR * S; e += C::T) {
      const int s = e / C::GR, r = e - s * C::GR;   // compile-time divisors
      const long J = J0 + r / NT;
      if (J < nJ) {
        const long R = J * NMU + g * NT + r % NT;
        at[(long)s * Mh + R] = G[r * C::GS + s];
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------------------
// Dispatch table
// ---------------------------------------------------------------------------------------------
template <int S, int NMU, int NT, int DMU, int BP, int U, int NP, int W, bool CT>
static int launch(const SoiFilterArgs *a, hipStream_t st)
{
  typedef FilterCfg<S, NMU, NT, DMU, BP, U, NP, W, CT> C;
  auto kern = k_filter<S, NMU, NT, DMU, BP, U, NP, W, CT>;
  const size_t lds = (size_t)(C::LDSN + S) * sizeof(double2);
  static bool once = false;
  if (!once) {
    (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    once = true;
  }
  const long ntiles = (a->nJ + C::TJ - 1) / C::TJ;
  // blocks per phase group: enough to fill the CUs, tiles split into contiguous ranges
  const int per_cu = (int)((160 * 1024) / lds) > 0 ? (int)((160 * 1024) / lds) : 1;
  long nb = 256L * (per_cu > 2 ? 2 : per_cu) / C::NG;
  if (nb < 1) nb = 1;
  if (nb > ntiles) nb = ntiles;
  const int tpb = (int)((ntiles + nb - 1) / nb);
  nb = (ntiles + tpb - 1) / tpb;
  hipLaunchKernelGGL(kern, dim3((unsigned)(nb * C::NG)), dim3(C::T), lds, st,
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
    *U = 16;
    *NT = 5;
    int half = (B + 1) / 2;
    *BP = half <= 32 ? 32 : half <= 36 ? 36 : half <= 40 ? 40 : 0;
    if (!*BP || B < 60) return -1;
    switch (S) {
      case 4: *W = 8; break;
      case 8: *W = 4; break;
      case 16: *W = 2; break;
      case 32: *W = 1; break;
      case 64: *W = 1; break;
      default: return -1;
    }
    return 0;
  }
  *NP = 4;
  int q = (B + 3) / 4;
  if (n_mu == 5 && d_mu == 4) {
    *U = 16;
    *BP = q <= 8 ? 8 : q <= 12 ? 12 : q <= 16 ? 16 : q <= 20 ? 20 : 0;
    if (!*BP) return -1;
    switch (S) {
      case 4: *NT = 5; *W = 4; break;
      case 8: *NT = 5; *W = 2; break;
      case 16: *NT = 5; *W = 1; break;
      case 32: *NT = 5; *W = 1; break;
      case 64: *NT = 1; *W = 1; break;
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
  return (size_t)((xn > gn ? xn : gn) + a->S) * sizeof(double2);
}

#define RT_CASE(SV, BPV, WV)                                                        \
  if (a->S == SV && a->BP == BPV) return launch<SV, 5, 5, 4, BPV, 16, 2, WV, false>(a, st);
#define CT54_CASE(SV, NTV, BPV, WV)                                                 \
  if (a->S == SV && a->BP == BPV) return launch<SV, 5, NTV, 4, BPV, 16, 4, WV, true>(a, st);
#define CT87_CASE(SV, NTV, BPV, WV)                                                 \
  if (a->S == SV && a->BP == BPV) return launch<SV, 8, NTV, 7, BPV, 8, 4, WV, true>(a, st);

extern "C" int soik_filter_sdft(const SoiFilterArgs *a, hipStream_t st)
{
  if (!a->ct) {
    RT_CASE(4, 32, 8) RT_CASE(4, 36, 8) RT_CASE(4, 40, 8)
    RT_CASE(8, 32, 4) RT_CASE(8, 36, 4) RT_CASE(8, 40, 4)
    RT_CASE(16, 32, 2) RT_CASE(16, 36, 2) RT_CASE(16, 40, 2)
    RT_CASE(32, 32, 1) RT_CASE(32, 36, 1) RT_CASE(32, 40, 1)
    RT_CASE(64, 32, 1) RT_CASE(64, 36, 1) RT_CASE(64, 40, 1)
  } else if (a->n_mu == 5) {
    CT54_CASE(4, 5, 8, 4) CT54_CASE(4, 5, 12, 4) CT54_CASE(4, 5, 16, 4) CT54_CASE(4, 5, 20, 4)
    CT54_CASE(8, 5, 8, 2) CT54_CASE(8, 5, 12, 2) CT54_CASE(8, 5, 16, 2) CT54_CASE(8, 5, 20, 2)
    CT54_CASE(16, 5, 8, 1) CT54_CASE(16, 5, 12, 1) CT54_CASE(16, 5, 16, 1) CT54_CASE(16, 5, 20, 1)
    CT54_CASE(32, 5, 8, 1) CT54_CASE(32, 5, 12, 1) CT54_CASE(32, 5, 16, 1) CT54_CASE(32, 5, 20, 1)
    CT54_CASE(64, 1, 8, 1) CT54_CASE(64, 1, 12, 1) CT54_CASE(64, 1, 16, 1) CT54_CASE(64, 1, 20, 1)
  } else {
    CT87_CASE(4, 8, 16, 2) CT87_CASE(4, 8, 20, 2)
    CT87_CASE(8, 8, 16, 1) CT87_CASE(8, 8, 20, 1)
    CT87_CASE(16, 8, 16, 1) CT87_CASE(16, 8, 20, 1)
    CT87_CASE(32, 4, 16, 1) CT87_CASE(32, 4, 20, 1)
    CT87_CASE(64, 2, 16, 1) CT87_CASE(64, 2, 20, 1)
  }
  return (int)hipErrorInvalidValue;
}
