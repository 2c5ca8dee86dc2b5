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
#include <stdlib.h>
#include <type_traits>
#include "soi_common.hpp"
#include "soi_kernels.h"

using namespace soi;

typedef __attribute__((address_space(3))) void lds_void_t;


template <int BP, bool CT, typename RT> struct Taps;
template <int BP, typename RT> struct Taps<BP, false, RT> { RT v[BP]; };
template <int BP, typename RT> struct Taps<BP, true, RT> { RT re[BP], im[BP]; };

// Fully unrolled accumulation: input row o of the thread's column feeds accumulator u with tap
// b = o - u*d whenever 0 <= b < BP (constant indices: taps and accumulators stay in registers).
template <int O, int UU, int U, int DMU, int BP, bool CT, typename C>
__device__ __forceinline__ void conv_taps(const C xv, const Taps<BP, CT, real_t<C>> &h, C (&acc)[U])
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
    conv_taps<O, UU + 1, U, DMU, BP, CT, C>(xv, h, acc);
  }
}

template <int O, int OEND, int S, int U, int DMU, int BP, bool CT, typename C>
__device__ __forceinline__ void conv_rows(const C *__restrict__ col, const Taps<BP, CT, real_t<C>> &h,
                                          C (&acc)[U])
{
  if constexpr (O < OEND) {
    const C xv = col[O * S];
    conv_taps<O, 0, U, DMU, BP, CT, C>(xv, h, acc);
    conv_rows<O + 1, OEND, S, U, DMU, BP, CT, C>(col, h, acc);
  }
}

template <int S, int NMU, int NT, int DMU, int BP, int U, int NP, int W, bool CT, typename C>
struct FilterCfg {
  static constexpr int EPL = 16 / (int)sizeof(C);    // values per 16-byte direct-to-LDS load
  static constexpr int T = S * NT * NP * W;          // threads
  static constexpr int NWAVES = T / 64;
  static constexpr int TJ = U * W;                   // row groups per tile
  static constexpr int ROWS = TJ * DMU + NP * BP - DMU;  // input rows per tile
  static constexpr int XN = (ROWS * S + 64 * EPL - 1) / (64 * EPL) * (64 * EPL);  // staged values
  static constexpr int NDMA = XN / (64 * EPL);       // wave-wide 16-byte loads per tile
  static constexpr int GR = TJ * NT;                 // output rows per tile (this block's phases)
  static constexpr int GS = S + 1;                   // padded G row stride
  static constexpr int GN = GR * GS;
  static constexpr int BUFN = (XN > GN ? XN : GN);   // one tile buffer (G reuses it)
  static constexpr int NG = NMU / NT;                // phase groups
  // two tile buffers (the next tile's rows stream in while this one computes) when they fit
  static constexpr bool DB = (2 * BUFN + S) * (int)sizeof(C) <= 160 * 1024;
  static constexpr int LDS_BYTES = ((DB ? 2 : 1) * BUFN + S) * (int)sizeof(C);
  // blocks resident per CU (LDS and registers), and the register budget that follows
  static constexpr int BPC = (LDS_BYTES <= 80 * 1024 && T <= 320) ? 2 : 1;
  static constexpr int MINW = (BPC * NWAVES + 3) / 4;
  // two real-tap slices whose partial sums meet by a lane exchange inside the wave: the slice
  // index is lane bit log2(S)
  // (complex64 only: for complex128 the lanes of the two phases that read the same rows are
  // worth more to the FMA block's LDS traffic, measured)
  static constexpr bool LSW = !CT && NP == 2 && S <= 32 && U % 2 == 0 && sizeof(C) == 8;
  static constexpr int NSC = (GR * S + T - 1) / T;   // stores per thread per tile (at most)
  static constexpr int NSC_MIN = (GR * S) / T;       // (at least)
};

constexpr int filter_threads(int S, int NT, int NP, int W) { return S * NT * NP * W; }

// s_waitcnt vmcnt(n) (expcnt and lgkmcnt left alone), through the builtin so the compiler's
// own wait tracking sees it
template <int n> __device__ __forceinline__ void wait_vmcnt()
{
  static_assert(n >= 0 && n < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((n & 15) | (7 << 4) | (15 << 8) | ((n >> 4) << 14));
}

// In-place S-point DFT of each of the GR rows of G (row r at G + r*GS, unit stride), Stockham
// passes of radix 8 then 4 / 2 with compile-time shapes; tw[m] = e^{-2 pi i m/S} in LDS.
// Butterfly g takes row g % GR: with the odd row stride GS, consecutive lanes then read and
// write distinct banks.
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

template <int S, int GR, int GS, int T, int P, int NS, typename C>
__device__ __forceinline__ void sdft_passes(C *G, const C *tw, int tid)
{
  constexpr int R = sdft_radix<S>(P);
  if constexpr (R != 0) {
    constexpr int NBF = S / R;                 // butterflies per row
    constexpr int TOT = GR * NBF;
    constexpr int NB = (TOT + T - 1) / T;
    C v[NB][R];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      const int g = tid + c * T;
      if (TOT % T == 0 || g < TOT) {
        const int q = g % GR, j = g / GR;  // rows fastest: consecutive lanes on distinct banks
        const C *src = G + q * GS + j;
#pragma unroll
        for (int k = 0; k < R; ++k) v[c][k] = src[k * NBF];
      }
    }
    lds_barrier();
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      const int g = tid + c * T;
      if (TOT % T == 0 || g < TOT) {
        const int q = g % GR, j = g / GR;  // rows fastest: consecutive lanes on distinct banks
        const int jm = j % NS;
        if constexpr (NS > 1) {
          constexpr int step = S / (NS * R);
#pragma unroll
          for (int k = 1; k < R; ++k) v[c][k] = cmul(v[c][k], tw[k * jm * step]);
        }
        dft_r<R>(v[c]);
        C *dst = G + q * GS + (j - jm) * R + jm;
#pragma unroll
        for (int k = 0; k < R; ++k) dst[k * NS] = v[c][k];
      }
    }
    lds_barrier();
    sdft_passes<S, GR, GS, T, P + 1, NS * R, C>(G, tw, tid);
  }
}

// Exchange of the two tap-slice partial sums held by lanes l and l ^ S (slice p = lane bit
// log2 S), complex64: a = acc[u], b = acc[u + U/2] on entry; on return the slice-0 lane holds
// acc0[u] + acc1[u] and the slice-1 lane acc0[u + U/2] + acc1[u + U/2].
template <int S>
__device__ __forceinline__ float2 slice_sum(float2 a, float2 b, int p)
{
  if constexpr (S == 32) {
    // v_permlane32_swap: lanes 32..63 of the first operand trade with lanes 0..31 of the second
    const auto re = __builtin_amdgcn_permlane32_swap(__float_as_uint(a.x), __float_as_uint(b.x),
                                                     false, false);
    const auto im = __builtin_amdgcn_permlane32_swap(__float_as_uint(a.y), __float_as_uint(b.y),
                                                     false, false);
    return float2{__uint_as_float(re[0]) + __uint_as_float(re[1]),
                  __uint_as_float(im[0]) + __uint_as_float(im[1])};
  } else {
    const float2 keep = p ? b : a, send = p ? a : b;
    return cadd(keep, float2{__shfl_xor(send.x, S), __shfl_xor(send.y, S)});
  }
}

// Stage the XN input values of a tile (contiguous from element g0, wrapping once at N) into
// dst with direct-to-LDS loads; the loading waves take 64-value chunks in turn.  A 10-wave
// block spreads 3/3/2/2 over the SIMDs (wave w on SIMD w % 4) and its FMA block ends last on
// SIMDs 0 and 1, so there only the four waves of SIMDs 2 and 3 issue loads; otherwise all do.
template <typename CF, typename C>
__device__ __forceinline__ void dma_rows(C *dst, const C *__restrict__ x, long g0, long N,
                                         int wave, int lane)
{
  constexpr int CH = 64 * CF::EPL;  // values per wave-wide load
  int slot = wave, nslot = CF::NWAVES;
  if constexpr (CF::NWAVES == 10) {
    if ((wave & 3) < 2) return;
    slot = (wave >> 2) * 2 + (wave & 3) - 2;
    nslot = 4;
  }
#pragma unroll
  for (int c = slot; c < CF::NDMA; c += nslot) {
    long gi = g0 + c * CH + lane * CF::EPL;  // N and g0 are multiples of EPL: no pair straddles N
    if (gi >= N) gi -= N;
    __builtin_amdgcn_global_load_lds((const void *)(x + gi), (lds_void_t *)(dst + c * CH), 16,
                                     0, 0);
  }
}

// Per-thread state that stays in registers for the whole kernel.
template <int BP, bool CT, typename C> struct FilterThread {
  Taps<BP, CT, real_t<C>> h;
  C pf;              // post phase
  int i, tl, p, w, g;
  int tid, wave, lane;
};

// One tile: prefetch the next tile's rows into nxt (when there is a next tile), FMA block from
// cur, tap-slice reduction + post phase + S-point DFT in cur, stores.  Barriers wait on LDS
// traffic only: the prefetch and the stores stay in flight.
template <typename CF, int S, int NMU, int NT, int DMU, int BP, int U, int NP, bool CT,
          typename C>
__device__ __forceinline__ void filter_tile(const FilterThread<BP, CT, C> &th, C *cur, C *nxt,
                                            const C *tw, const C *__restrict__ x,
                                            const C *__restrict__ xn, C *__restrict__ at,
                                            long tile, long tnext, bool more, long N, long Mh,
                                            long nJ)
{
  // x: this tile's input; xn, tnext: the next tile's input and index (prefetched when more)
  const long J0 = tile * CF::TJ;
  if constexpr (CF::DB) {
    if (more) dma_rows<CF, C>(nxt, xn, tnext * CF::TJ * DMU * S, N, th.wave, th.lane);
  } else {
    dma_rows<CF, C>(cur, x, tile * CF::TJ * DMU * S, N, th.wave, th.lane);
    wait_vmcnt<0>();
    lds_barrier();
  }

  // ---- FMA block
  C acc[U];
#pragma unroll
  for (int u = 0; u < U; ++u) acc[u] = C{0, 0};
  conv_rows<0, (U - 1) * DMU + BP, S, U, DMU, BP, CT, C>(cur + (th.w * U * DMU + th.p * BP) * S + th.i,
                                                      th.h, acc);
  lds_barrier();

  // ---- tap-slice reduction and post phase, into G = cur
  C *G = cur;
  const int gbase = (th.w * U) * NT + th.tl;
  if constexpr (CF::LSW) {
    // lane exchange: each slice finishes half of the U outputs (acc0 + acc1 in both)
#pragma unroll
    for (int u = 0; u < U / 2; ++u) {
      const C a = cmul(slice_sum<S>(acc[u], acc[u + U / 2], th.p), th.pf);
      G[(gbase + (u + th.p * (U / 2)) * NT) * CF::GS + th.i] = a;
    }
    lds_barrier();
  } else
#pragma unroll
  for (int q = NP - 1; q >= 0; --q) {
    if (th.p == q) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        C a = acc[u];
        C *gp = G + (gbase + u * NT) * CF::GS + th.i;
        if (q != NP - 1) a = cadd(a, *gp);
        if (q == 0 && !CT) a = cmul(a, th.pf);
        *gp = a;
      }
    }
    lds_barrier();
  }

  // ---- S-point DFT of each row
  sdft_passes<S, CF::GR, CF::GS, CF::T, 0, 1, C>(G, tw, th.tid);

  // ---- store segment-major: R = J*n_mu + theta (r fastest: runs of consecutive R)
#pragma unroll
  for (int c = 0; c < CF::NSC; ++c) {
    const int e = th.tid + c * CF::T;
    if ((CF::GR * S) % CF::T == 0 || e < CF::GR * S) {
      const int s = e / CF::GR, r = e % CF::GR;
      const long J = J0 + r / NT;
      if (J < nJ) {
        const long R = J * NMU + th.g * NT + r % NT;
        store_nt(at + (long)s * Mh + R, G[r * CF::GS + s]);
      }
    }
  }
  if constexpr (CF::DB) {
    // this wave's prefetch loads are complete once at most its stores are outstanding (the
    // stores were issued after them; a whole tile stores every row, the last partial tile of a
    // transform fewer)
    asm volatile("" ::: "memory");
    if (more) {
      if (J0 + CF::TJ <= nJ) wait_vmcnt<CF::NSC_MIN>();
      else wait_vmcnt<0>();
    }
  }
  lds_barrier();
}

// Input and output of member m of a batched launch (the pointers passed directly for one
// transform)
template <typename C>
__device__ __forceinline__ const C *member_in(const C *x, const SoiBatch &b, long m)
{
  return b.x ? reinterpret_cast<const C *>(b.x[m]) : x;
}
template <typename C> __device__ __forceinline__ C *member_out(C *at, const SoiBatch &b, long m)
{
  return b.at ? reinterpret_cast<C *>(b.at[m]) : at;
}

template <int S, int NMU, int NT, int DMU, int BP, int U, int NP, int W, bool CT, typename C,
          bool BATCHED>
__global__ __launch_bounds__(filter_threads(S, NT, NP, W),
                             (FilterCfg<S, NMU, NT, DMU, BP, U, NP, W, CT, C>::MINW)) void k_filter(
    const C *__restrict__ x, C *__restrict__ at, const real_t<C> *__restrict__ taps,
    const C *__restrict__ post, const C *__restrict__ twS, long M, long Mh, long nJ,
    int tiles_per_block, SoiBatch b)
{
  typedef FilterCfg<S, NMU, NT, DMU, BP, U, NP, W, CT, C> CF;
  typedef real_t<C> RT;
  __shared__ C bufA[CF::BUFN];
  __shared__ C bufB[CF::DB ? CF::BUFN : 1];
  __shared__ C tw[S];
  FilterThread<BP, CT, C> th;
  th.tid = threadIdx.x;
  th.wave = th.tid >> 6;
  th.lane = th.tid & 63;
  th.i = th.tid % S;
  int rest = th.tid / S;
  if constexpr (CF::LSW) {  // (i, p, theta, w), p a lane bit
    th.p = rest % NP;
    rest /= NP;
    th.tl = rest % NT;
    th.w = rest / NT;
  } else {                  // (i, theta, p, w)
    th.tl = rest % NT;
    rest /= NT;
    th.p = rest % NP;
    th.w = rest / NP;
  }

  // phase group: the slow index of the block id so blocks sharing input rows are 8-apart
  // (same XCD under round-robin dispatch) and walk the same tiles at the same time
  const int nblk_g = gridDim.x / CF::NG;
  th.g = blockIdx.x / nblk_g;
  const int bg = blockIdx.x - th.g * nblk_g;
  const int theta = th.g * NT + th.tl;
  {
    const long base = ((long)theta * S + th.i) * (NP * BP) + th.p * BP;
    if constexpr (CT) {
      const C *tp = reinterpret_cast<const C *>(taps) + base;
#pragma unroll
      for (int b = 0; b < BP; ++b) {
        C v = tp[b];
        th.h.re[b] = v.x;
        th.h.im[b] = v.y;
      }
    } else {
      const RT *tp = taps + base;
#pragma unroll
      for (int b = 0; b < BP; ++b) th.h.v[b] = tp[b];
    }
  }
  th.pf = post[theta * S + th.i];
  for (int e = th.tid; e < S; e += CF::T) tw[e] = twS[e];

  const long ntiles = (nJ + CF::TJ - 1) / CF::TJ;
  const long N = M * S;
  if constexpr (!BATCHED) {
    const long t0 = (long)bg * tiles_per_block;
    const long t1 = min(ntiles, t0 + tiles_per_block);
    if constexpr (CF::DB) {
      if (t0 < t1) dma_rows<CF, C>(bufA, x, t0 * CF::TJ * DMU * S, N, th.wave, th.lane);
      wait_vmcnt<0>();
      lds_barrier();
      for (long tile = t0; tile < t1; ++tile) {
        const bool odd = (tile - t0) & 1;
        filter_tile<CF, S, NMU, NT, DMU, BP, U, NP, CT, C>(th, odd ? bufB : bufA, odd ? bufA : bufB,
                                                        tw, x, x, at, tile, tile + 1, tile + 1 < t1,
                                                        N, Mh, nJ);
      }
    } else {
      lds_barrier();
      for (long tile = t0; tile < t1; ++tile)
        filter_tile<CF, S, NMU, NT, DMU, BP, U, NP, CT, C>(th, bufA, bufA, tw, x, x, at, tile, tile,
                                                        false, N, Mh, nJ);
    }
  } else {
    // the tiles of all members form one range (member-major), split into contiguous block
    // ranges: a block keeps its taps across the members of its range
    const long total = ntiles * b.n;
    const long t0 = (long)bg * tiles_per_block;
    const long t1 = min(total, t0 + tiles_per_block);
    if (t0 >= t1) return;
    long m = t0 / ntiles, tile = t0 - m * ntiles;
    const C *xc = member_in(x, b, m);
    C *ac = member_out(at, b, m);
    if constexpr (CF::DB) {
      dma_rows<CF, C>(bufA, xc, tile * CF::TJ * DMU * S, N, th.wave, th.lane);
      wait_vmcnt<0>();
    }
    lds_barrier();
    for (long f = t0; f < t1; ++f) {
      const bool odd = (f - t0) & 1, more = f + 1 < t1;
      long tn = tile + 1, mn = m;
      if (tn == ntiles) tn = 0, ++mn;
      const C *xn = mn == m || !more ? xc : member_in(x, b, mn);
      if constexpr (CF::DB)
        filter_tile<CF, S, NMU, NT, DMU, BP, U, NP, CT, C>(th, odd ? bufB : bufA, odd ? bufA : bufB,
                                                        tw, xc, xn, ac, tile, tn, more, N, Mh, nJ);
      else
        filter_tile<CF, S, NMU, NT, DMU, BP, U, NP, CT, C>(th, bufA, bufA, tw, xc, xc, ac, tile, tile,
                                                        false, N, Mh, nJ);
      if (!more) break;
      tile = tn;
      if (mn != m) {
        m = mn;
        xc = xn;
        ac = member_out(at, b, m);
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Complex64 real-tap filter on MFMA with split-bf16 operands: every FP32 value is hi + mid + lo
// bf16 pieces (24 significant bits), and six of the nine piece products (hh, hm, mh, hl, lh, mm;
// the dropped ones are below 2^-24 relative) accumulate in FP32.
//
// For lane i, the outputs of row groups J = J0 + 3m + D (m < 8, D < 3) and phase theta form one
// 16 x 16 x K real product per tile and lane:
//   C[2m + part][theta + 5D] = sum_k x_i[4 J0 + 12 m + k].part * t_i[theta][k - 4D]
// with t_i[theta][b] the real taps (zero outside [0, BPAD)) and K = BPAD + 8 <= 96 covered by
// three 32-deep chunks (v_mfma_f32_16x16x32_bf16; every accumulation chain stays on this one
// opcode).
// So R = 5 J + theta = 5 J0 + 15 m + n for column n = theta + 5D < 15.  A wave owns LPW lanes i
// and keeps their B operands (the tap pieces) in registers for the whole kernel; the A operands
// come from an LDS copy of the tile's input rows transposed to one run per (lane, piece, part).
//
// Per tile: the raw rows of the next tile arrive by direct-to-LDS loads while this one computes;
// the split pass turns raw rows into bf16 pieces; MFMA; post-phase multiply into G; the S-point
// DFT of each row of G; coalesced stores of 120 consecutive R per segment.
// ---------------------------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int S> struct MfmaCfg {
  static constexpr int EPL = 2;                 // complex64 values per 16-byte direct-to-LDS load
  static constexpr int TJ = 24;                 // row groups per tile: 8 MFMA rows x 3 shifts
  static constexpr int NCH = 3;                 // 32-deep K chunks
  static constexpr int K = 32 * NCH;
  static constexpr int NRG = (12 * 7 + K + 7) / 8;  // 8-row groups staged per tile
  static constexpr int ROWS = NRG * 8;
  static constexpr int RP = ROWS;               // bf16 pitch of one (lane, piece, part) run
  static constexpr int LPW = S >= 32 ? 4 : S < 2 ? S : 2;  // lanes per wave (<= 8 waves)
  static constexpr int NWAVES = S / LPW;
  static constexpr int T = NWAVES * 64;
  static constexpr int XN = (ROWS * S + 64 * EPL - 1) / (64 * EPL) * (64 * EPL);  // raw values
  static constexpr int NDMA = XN / (64 * EPL);  // wave-wide 16-byte loads per tile
  static constexpr int GR = TJ * 5;             // output rows R per tile
  static constexpr int GS = S + 1;
  static constexpr int NSC = (GR * S + T - 1) / T;   // stores per thread per tile (at most)
  static constexpr int NSC_MIN = (GR * S) / T;        // (at least)
};

// hi + mid + lo bf16 split of an FP32 value (each residual is exact in FP32)
__device__ __forceinline__ void split3(float x, __bf16 &h, __bf16 &m, __bf16 &l)
{
  h = (__bf16)x;
  const float r1 = x - (float)h;
  m = (__bf16)r1;
  l = (__bf16)(r1 - (float)m);
}

template <int S>
__global__ __launch_bounds__(MfmaCfg<S>::T, MfmaCfg<S>::T / 256) void k_filter_mfma(
    const float2 *__restrict__ x, float2 *__restrict__ at, const bf16x8 *__restrict__ wt,
    const float2 *__restrict__ post, const float2 *__restrict__ twS, long M, long Mh, long nJ,
    int tiles_per_block, SoiBatch b)
{
  typedef MfmaCfg<S> CF;
  x = batch_ptr(x, b.x);
  at = batch_ptr(at, b.at);
  constexpr int LPW = CF::LPW, NCH = CF::NCH, RP = CF::RP;
  __shared__ __align__(16) float2 raw[CF::XN];
  __shared__ __align__(16) __bf16 As[S * 6 * RP];
  __shared__ __align__(16) float2 G[CF::GR * CF::GS];
  __shared__ __align__(16) float2 tw[S];
  __shared__ __align__(16) float2 pfl[5 * S];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int n = lane & 15, g = lane >> 4;

  // B operands (tap pieces) of this wave's lanes
  bf16x8 wb[LPW][NCH][3];
#pragma unroll
  for (int li = 0; li < LPW; ++li) {
    const int i = wave * LPW + li;
#pragma unroll
    for (int kc = 0; kc < NCH; ++kc)
#pragma unroll
      for (int p = 0; p < 3; ++p) wb[li][kc][p] = wt[((i * NCH + kc) * 3 + p) * 64 + lane];
  }
  for (int e = tid; e < S; e += CF::T) tw[e] = twS[e];
  for (int e = tid; e < 5 * S; e += CF::T) pfl[e] = post[e];

  const long ntiles = (nJ + CF::TJ - 1) / CF::TJ;
  const long t0 = (long)blockIdx.x * tiles_per_block;
  const long t1 = min(ntiles, t0 + tiles_per_block);
  const long N = M * S;
  if (t0 >= t1) return;

  // raw rows of a tile -> bf16 pieces, As[((i*3 + p)*2 + part)*RP + row]
  auto split_tile = [&]() {
    for (int task = tid; task < CF::NRG * S; task += CF::T) {
      const int i = task % S, rg = task / S;
#pragma unroll
      for (int part = 0; part < 2; ++part) {
        bf16x8 pc[3];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          const float2 v = raw[(rg * 8 + r) * S + i];
          __bf16 h, m, l;
          split3(part ? v.y : v.x, h, m, l);
          pc[0][r] = h, pc[1][r] = m, pc[2][r] = l;
        }
#pragma unroll
        for (int p = 0; p < 3; ++p)
          *reinterpret_cast<bf16x8 *>(As + ((i * 3 + p) * 2 + part) * RP + rg * 8) = pc[p];
      }
    }
  };

  // the raw rows of tile t + 1 stream in from the end of tile t - 1's pieces split to the end
  // of tile t's MFMA block
  dma_rows<CF, float2>(raw, x, t0 * CF::TJ * 4 * S, N, wave, lane);
  wait_vmcnt<0>();
  lds_barrier();
  split_tile();
  lds_barrier();
  if (t0 + 1 < t1) dma_rows<CF, float2>(raw, x, (t0 + 1) * CF::TJ * 4 * S, N, wave, lane);

  // A operand row of this lane: (m, part) = (r >> 1, r & 1), r = lane % 16
  const int am = n >> 1, apart = n & 1;
  for (long tile = t0; tile < t1; ++tile) {
    const bool more = tile + 1 < t1;

    // ---- MFMA: C_li = sum over K chunks of the six piece products
    f32x4 acc[LPW];
#pragma unroll
    for (int li = 0; li < LPW; ++li) acc[li] = f32x4{0, 0, 0, 0};
#pragma unroll
    for (int kc = 0; kc < NCH; ++kc) {
#pragma unroll
      for (int li = 0; li < LPW; ++li) {
        const int i = wave * LPW + li;
        bf16x8 a[3];
#pragma unroll
        for (int p = 0; p < 3; ++p) {  // K offset 8g: 8 rows at an 8-byte aligned address
          const __bf16 *src = As + ((i * 3 + p) * 2 + apart) * RP + 12 * am + 32 * kc + 8 * g;
          const bf16x4 lo4 = *reinterpret_cast<const bf16x4 *>(src);
          const bf16x4 hi4 = *reinterpret_cast<const bf16x4 *>(src + 4);
          a[p] = bf16x8{lo4[0], lo4[1], lo4[2], lo4[3], hi4[0], hi4[1], hi4[2], hi4[3]};
        }
        f32x4 c = acc[li];
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], wb[li][kc][1], c, 0, 0, 0);  // mm
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], wb[li][kc][0], c, 0, 0, 0);  // lh
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], wb[li][kc][2], c, 0, 0, 0);  // hl
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], wb[li][kc][0], c, 0, 0, 0);  // mh
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], wb[li][kc][1], c, 0, 0, 0);  // hm
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], wb[li][kc][0], c, 0, 0, 0);  // hh
        acc[li] = c;
      }
    }
    lds_barrier();  // every wave is done reading As

    // ---- post phase into G: accumulator j of this lane is part j & 1 of row m = 2g + (j >> 1)
    if (n < 15) {
#pragma unroll
      for (int li = 0; li < LPW; ++li) {
        const int i = wave * LPW + li;
        const float2 pf = pfl[(n % 5) * S + i];
#pragma unroll
        for (int h = 0; h < 2; ++h)
          G[(15 * (2 * g + h) + n) * CF::GS + i] = cmul(float2{acc[li][2 * h], acc[li][2 * h + 1]}, pf);
      }
    }
    if (more) {
      // the next tile's raw rows are complete once at most this wave's stores of the previous
      // tile are outstanding (they were issued after the rows; none before the first tile)
      asm volatile("" ::: "memory");
      if (tile == t0) wait_vmcnt<0>();
      else wait_vmcnt<CF::NSC_MIN>();
      lds_barrier();
      split_tile();
    }
    lds_barrier();  // publishes G and the next tile's pieces; raw is free again
    if (tile + 2 < t1) dma_rows<CF, float2>(raw, x, (tile + 2) * CF::TJ * 4 * S, N, wave, lane);

    // ---- S-point DFT of each row
    sdft_passes<S, CF::GR, CF::GS, CF::T, 0, 1, float2>(G, tw, tid);

    // ---- store segment-major: R = 5 J0 + r (runs of GR consecutive R per segment)
    const long R0 = tile * CF::TJ * 5;
#pragma unroll
    for (int c = 0; c < CF::NSC; ++c) {
      const int e = tid + c * CF::T;
      if ((CF::GR * S) % CF::T == 0 || e < CF::GR * S) {
        const int s = e / CF::GR, r = e % CF::GR;
        if (tile * CF::TJ + r / 5 < nJ) store_nt(at + (long)s * Mh + R0 + r, G[r * CF::GS + s]);
      }
    }
  }
}

// B operand table: wt[((i*NCH + ch)*3 + p)*64 + lane] holds bf16 piece p (hi, mid, lo) of
// t_i[theta][k - 4D] for k = 32 ch + 8 (lane/16) + j (j < 8) and n = lane % 16 = theta + 5D
// (zero for n = 15 or a tap index outside [0, BPAD)); t from the FP64 real-tap table
// taps[(theta*S + i)*BPAD + b].
__global__ void k_filter_mfma_table(const double *__restrict__ taps, unsigned short *__restrict__ wt,
                                    int S, int BPAD, int NCH)
{
  const long total = (long)S * NCH * 64 * 8;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int j = (int)(e % 8), lane = (int)(e / 8 % 64), ch = (int)(e / 512 % NCH), i = (int)(e / (512L * NCH));
    const int n = lane % 16;
    const int k = 32 * ch + 8 * (lane / 16) + j;
    const int theta = n % 5, D = n / 5, b = k - 4 * D;
    const bool on = n < 15 && b >= 0 && b < BPAD;
    const double t = on ? taps[((long)theta * S + i) * BPAD + b] : 0.0;
    const __bf16 h = (__bf16)(float)t;
    const double r1 = t - (double)(float)h;
    const __bf16 m = (__bf16)(float)r1;
    const __bf16 l = (__bf16)(float)(r1 - (double)(float)m);
    const long base = (((long)i * NCH + ch) * 3) * 64 * 8 + lane * 8 + j;
    wt[base] = __builtin_bit_cast(unsigned short, h);
    wt[base + 512] = __builtin_bit_cast(unsigned short, m);
    wt[base + 1024] = __builtin_bit_cast(unsigned short, l);
  }
}

// Blocks per transform of a complex128 MFMA filter launch.  One transform: 256 (one per CU, each
// walking a contiguous range of tiles).  A batched launch (blockIdx.y = member) shares about two
// rounds of blocks over all its members, so that each block's prologue (its B operands: up to
// 160 KB of tap table at S = 16) is spent over many tiles also when the members are small.
// Every whole tile goes through the same loop body, so a member's results do not depend on how
// its tiles are split over blocks (the complex64 kernels write each block's last tile through a
// separate code copy and keep the single-transform split, so that batched results stay
// bit-identical to eager ones).  SOI_FILTER_BATCH_BLOCKS sets the total (0: 256 per member).
static long mfma_blocks_per_member(const SoiFilterArgs *a)
{
  static const long total = [] {
    const char *e = getenv("SOI_FILTER_BATCH_BLOCKS");
    return e ? atol(e) : 512L;
  }();
  if (a->b.n <= 1 || total <= 0) return 256;
  const long nb = (total + a->b.n - 1) / a->b.n;
  return nb < 1 ? 1 : nb;
}

template <int S> static int launch_mfma(const SoiFilterArgs *a, hipStream_t st)
{
  typedef MfmaCfg<S> CF;
  if ((long)CF::XN > a->M * S) return (int)hipErrorInvalidValue;  // tile window wraps at most once
  const long ntiles = (a->nJ + CF::TJ - 1) / CF::TJ;
  long nb = 256;
  if (nb > ntiles) nb = ntiles;
  const int tpb = (int)((ntiles + nb - 1) / nb);
  nb = (ntiles + tpb - 1) / tpb;
  hipLaunchKernelGGL(k_filter_mfma<S>, dim3((unsigned)nb, a->b.n > 0 ? a->b.n : 1), dim3(CF::T), 0,
                     st, (const float2 *)a->x, (float2 *)a->at, (const bf16x8 *)a->wt,
                     (const float2 *)a->post, (const float2 *)a->twS, a->M, a->Mh, a->nJ, tpb,
                     a->b);
  return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Complex64 real-tap filter on MFMA with FP32 operands (v_mfma_f32_16x16x4_f32: exact FP32
// products, FP32 accumulation).  The GEMM of each lane i is the one of the split-bf16 form,
//   C[2m + part][theta + 5D] = sum_k x_i[4 J0 + 12 m + k].part * t_i[theta][k - 4D],
// with K = 80 >= BPAD + 8 covered by 20 four-deep steps.  The A operand of step ks in lane l is
// row (m, part) = ((l % 16) >> 1, l & 1), k = l / 16, read straight from the tile rows as the
// direct-to-LDS loads left them: 1 KB chunks with a 16-byte gap after each, so the 16 (m, part)
// rows of one read fall on distinct banks.  Result layout (v_mfma_f32_16x16x4_f32): D[row][col]
// in lane col + 16 (row / 4), register row % 4, so lane col + 16g holds m = 2g (registers 0, 1:
// real, imaginary) and m = 2g + 1 (registers 2, 3).
//
// Wave w of the NW = S / LPW waves owns the lanes i = NW li + w; the S-point DFT of each output
// row splits as in the complex128 MFMA filter: LPW-point DFTs in registers, one exchange
// through LDS, an NW-point DFT per (R, s1) whose results go straight to HBM.
//
// Two row buffers alternate: the loads of tile t + 2 fill tile t's buffer once its MFMA block is
// done, and complete under tile t + 1.  The tile loop is unrolled by two so that each half names
// its buffer statically: the compiler's wait tracking then sees that the loads in flight write
// the other buffer, and the MFMA reads wait for nothing.  Whole tiles issue the same stores in
// every wave, unconditionally; the transform's partial last tile runs after the loop.  The B
// operands (20 per lane i) stay in registers.
// ---------------------------------------------------------------------------------------------
typedef float f32x2g __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) f32x2g gf32x2;

__device__ __forceinline__ void store_nt_global(gf32x2 *p, float2 v)
{
  __builtin_nontemporal_store(f32x2g{v.x, v.y}, p);
}

template <int S, int LPW> struct Mfma32Cfg {
  static constexpr int EPL = 2;                 // complex64 values per 16-byte direct-to-LDS load
  static constexpr int TJ = 24;                 // row groups per tile: 8 MFMA rows x 3 shifts
  static constexpr int NK = 20;                 // four-deep K steps
  static constexpr int K = 4 * NK;
  static constexpr int ROWS = 12 * 7 + K;       // input rows read per tile
  static constexpr int NWAVES = S / LPW;
  static constexpr int T = NWAVES * 64;
  static constexpr int CH = 64 * EPL;           // values per wave-wide load (one chunk)
  static constexpr int CP = CH + 2;             // chunk pitch in LDS: a 16-byte gap
  static constexpr int NDMA = (ROWS * S + CH - 1) / CH;  // chunks per tile
  static constexpr int RAWN = NDMA * CP;
  static constexpr int GR = TJ * 5;             // output rows R per tile
  static constexpr int GP = S + 1;              // exchange row pitch: [R][s1 * NWAVES + w]
  static constexpr int NTASK = GR * LPW;
  static constexpr int NC = (NTASK + T - 1) / T;
  static constexpr int DUPD = NC * T - NTASK;
  static constexpr bool DUP = DUPD < 64;
  static_assert(LPW == 2 || LPW == 4 || LPW == 8, "in-register DFT size");
};

constexpr int mfma32_threads(int S, int LPW) { return S / LPW * 64; }

template <int S, int LPW>
__global__ __launch_bounds__(mfma32_threads(S, LPW), mfma32_threads(S, LPW) / 256) void k_filter_mfma32(
    const float2 *__restrict__ x, float2 *__restrict__ at, const float *__restrict__ wt,
    const float2 *__restrict__ post, const float2 *__restrict__ twS, long M, long Mh, long nJ,
    int tiles_per_block, SoiBatch b)
{
  typedef Mfma32Cfg<S, LPW> CF;
  x = batch_ptr(x, b.x);
  at = batch_ptr(at, b.at);
  constexpr int NK = CF::NK, CH = CF::CH, CP = CF::CP, NW = CF::NWAVES, GP = CF::GP;
  gf32x2 *const atg = (gf32x2 *)at;
  __shared__ __align__(16) float2 raw0[CF::RAWN];
  __shared__ __align__(16) float2 raw1[CF::RAWN];
  __shared__ __align__(16) float2 G[CF::GR * GP];
  __shared__ __align__(16) float2 tw[S];
  __shared__ __align__(16) float2 pfl[5 * S];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int n = lane & 15, g = lane >> 4;

  float wb[LPW][NK];  // B operands of this wave's lanes i = NW li + wave
#pragma unroll
  for (int li = 0; li < LPW; ++li)
#pragma unroll
    for (int ks = 0; ks < NK; ++ks) wb[li][ks] = wt[((li * NW + wave) * NK + ks) * 64 + lane];
  for (int e = tid; e < S; e += CF::T) tw[e] = twS[e];
  for (int e = tid; e < 5 * S; e += CF::T) pfl[e] = post[e];

  const long ntiles = (nJ + CF::TJ - 1) / CF::TJ;
  const long t0 = (long)blockIdx.x * tiles_per_block;
  const long t1 = min(ntiles, t0 + tiles_per_block);
  const long N = M * S;
  if (t0 >= t1) return;

  // the NDMA chunks of a tile's rows (contiguous from element g0, wrapping once at N)
  auto dma = [&](float2 *dst, long g0) {
#pragma unroll
    for (int c = wave; c < CF::NDMA; c += NW) {
      long gi = g0 + c * CH + lane * CF::EPL;  // N and g0 are multiples of EPL
      if (gi >= N) gi -= N;
      __builtin_amdgcn_global_load_lds((const void *)(x + gi), (lds_void_t *)(dst + c * CP), 16, 0, 0);
    }
  };

  dma(raw0, t0 * CF::TJ * 4 * S);
  wait_vmcnt<0>();
  lds_barrier();
  if (t0 + 1 < t1) dma(raw1, (t0 + 1) * CF::TJ * 4 * S);

  // A operand of lane (n, g): row (m, part) = (n >> 1, n & 1), k = g; value index within the
  // tile (12 m + 4 ks + g) * S + i with i = NW li + wave
  const int am = n >> 1, apart = n & 1;
  const int v0 = (12 * am + g) * S + wave;

  // One tile from buffer cur (B = 0: raw0, 1: raw1).  PART: the transform's partial last tile.
  auto body = [&](long tile, auto buf_tag, auto part_tag) {
    constexpr int BUF = decltype(buf_tag)::value;
    constexpr bool PART = decltype(part_tag)::value;
    float2 *cur = BUF ? raw1 : raw0;
    const float *curf = reinterpret_cast<const float *>(cur);
    int tv = tid;
    asm volatile("" : "+v"(tv));
    const int ln = tv & 63, nv = ln & 15, gv = ln >> 4;

    // ---- MFMA: one accumulation chain per lane i, all on v_mfma_f32_16x16x4_f32
    f32x4 acc[LPW];
#pragma unroll
    for (int li = 0; li < LPW; ++li) acc[li] = f32x4{0, 0, 0, 0};
    // A operand offsets from an opaque copy of the lane's first value index: computed once per
    // kernel they would be loop-invariant and hoisted (80 live registers).  When a K step is a
    // whole number of chunks (4S a multiple of CH) and the lanes i of a wave share a chunk, a
    // read is the lane's base plus a compile-time offset.
    int vb = v0;
    asm volatile("" : "+v"(vb));
    constexpr bool FAST = (4 * S) % CH == 0;
    const int abase = 2 * ((vb / CH) * CP + vb % CH) + apart;
#pragma unroll
    for (int ks = 0; ks < NK; ++ks) {
#pragma unroll
      for (int li = 0; li < LPW; ++li) {
        float a;
        if constexpr (FAST) {
          a = curf[abase + 2 * (ks * (4 * S / CH) * CP + li * NW)];
        } else {
          const int e = vb + 4 * ks * S + li * NW;
          a = curf[2 * ((e / CH) * CP + e % CH) + apart];
        }
        acc[li] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, wb[li][ks], acc[li], 0, 0, 0);
      }
    }
    lds_barrier();  // every wave is done reading cur and the previous tile's G

    // ---- post phase and the LPW-point DFTs over this wave's lanes i, twiddled by w_S^{w s1},
    // into the exchange rows G[R][s1 * NW + w]; column nv = theta + 5D (15 is unused)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      float2 v[LPW];
#pragma unroll
      for (int li = 0; li < LPW; ++li)
        v[li] = cmul(float2{acc[li][2 * h], acc[li][2 * h + 1]}, pfl[(nv % 5) * S + li * NW + wave]);
      dft_r<LPW>(v);
      if (nv < 15) {
        float2 *gr = G + (15 * (2 * gv + h) + nv) * GP + wave;
        gr[0] = v[0];
#pragma unroll
        for (int s1 = 1; s1 < LPW; ++s1) gr[s1 * NW] = cmul(v[s1], tw[(wave * s1) % S]);
      }
    }
    if constexpr (!PART) {
      // the next tile's rows are complete once at most this wave's stores of the previous tile
      // are outstanding (issued after those rows; none before the block's first tile).
      // Unconditional in whole tiles (harmless when no tile follows), so that every path back to
      // the loop head has waited.
      asm volatile("" ::: "memory");
      if (tile == t0) wait_vmcnt<0>();
      else if constexpr (CF::DUP) wait_vmcnt<CF::NC * NW>();
      else wait_vmcnt<0>();
    }
    lds_barrier();  // publishes G and the next tile's rows; cur is free
    if (!PART && tile + 2 < t1) dma(cur, (tile + 2) * CF::TJ * 4 * S);

    // ---- NW-point DFT over the waves for each (R, s1), stored segment-major
    const long R0 = tile * CF::TJ * 5;
#pragma unroll
    for (int c = 0; c < CF::NC; ++c) {
      int e = tv + c * CF::T;
      if constexpr (CF::DUP) {
        if (e >= CF::NTASK) e -= CF::DUPD;
      }
      if (CF::DUP || CF::NTASK % CF::T == 0 || e < CF::NTASK) {
        const int r = e % CF::GR, s1 = e / CF::GR;
        float2 u[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) u[w] = G[r * GP + s1 * NW + w];
        dft_r<NW>(u);
        if (!PART || tile * CF::TJ + r / 5 < nJ) {
#pragma unroll
          for (int s2 = 0; s2 < NW; ++s2)
            store_nt_global(atg + (long)(s1 + LPW * s2) * Mh + R0 + r, u[s2]);
        }
      }
    }
  };

  // whole tiles, two per iteration (raw0 then raw1); the partial last tile after them
  const long tfull = min(t1, nJ / CF::TJ);
  long tile = t0;
  for (; tile + 1 < tfull; tile += 2) {
    body(tile, std::integral_constant<int, 0>{}, std::false_type{});
    body(tile + 1, std::integral_constant<int, 1>{}, std::false_type{});
  }
  if (tile < tfull) {
    body(tile, std::integral_constant<int, 0>{}, std::false_type{});
    ++tile;
  }
  if (tile < t1) {
    if ((tile - t0) & 1) body(tile, std::integral_constant<int, 1>{}, std::true_type{});
    else body(tile, std::integral_constant<int, 0>{}, std::true_type{});
  }
}

// ---------------------------------------------------------------------------------------------
// The same complex64 filter software-pipelined across tiles: the epilogue of tile t (post phase,
// LPW-point DFTs, exchange, NW-point DFT and stores) shares its two barrier intervals with the
// MFMA chain of tile t + 1 (ten K steps in each).  Waves 0-3 and 4-7 (one of each on every SIMD)
// take the two halves of an interval in opposite orders, so that one wave's VALU and LDS work
// issues while the other wave's MFMAs occupy the matrix core.  The accumulators of tiles t and
// t + 1 are both live (2 x 4 x LPW registers).
//
// Step t (tile t's accumulators in acc, tile t + 1's rows in buffer NB, the other buffer free):
//   loads of tile t + 2 into the free buffer
//   | MFMA ks 0-9 of tile t + 1   and   post + LPW-point DFTs + exchange writes of tile t |
//   barrier (exchange rows of tile t published)
//   | MFMA ks 10-19 of tile t + 1  and   exchange reads, NW-point DFTs, stores of tile t  |
//   wait for the loads of tile t + 2, barrier (buffers and exchange rows free)
// ---------------------------------------------------------------------------------------------
template <int S, int LPW>
__global__ __launch_bounds__(mfma32_threads(S, LPW), mfma32_threads(S, LPW) / 256) void k_filter_mfma32p(
    const float2 *__restrict__ x, float2 *__restrict__ at, const float *__restrict__ wt,
    const float2 *__restrict__ post, const float2 *__restrict__ twS, long M, long Mh, long nJ,
    int tiles_per_block, SoiBatch b)
{
  typedef Mfma32Cfg<S, LPW> CF;
  x = batch_ptr(x, b.x);
  at = batch_ptr(at, b.at);
  constexpr int NK = CF::NK, CH = CF::CH, CP = CF::CP, NW = CF::NWAVES, GP = CF::GP;
  constexpr int KH = NK / 2;
  static_assert(NK % 2 == 0, "two halves of the K steps");
  gf32x2 *const atg = (gf32x2 *)at;
  __shared__ __align__(16) float2 raw0[CF::RAWN];
  __shared__ __align__(16) float2 raw1[CF::RAWN];
  __shared__ __align__(16) float2 G[CF::GR * GP];
  __shared__ __align__(16) float2 tw[S];
  __shared__ __align__(16) float2 pfl[5 * S];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int n = lane & 15, g = lane >> 4;
  // the wave group that runs its epilogue first (wave-uniform: a scalar branch)
  const bool late = (__builtin_amdgcn_readfirstlane(wave) >> 2) & 1;

  float wb[LPW][NK];  // B operands of this wave's lanes i = NW li + wave
#pragma unroll
  for (int li = 0; li < LPW; ++li)
#pragma unroll
    for (int ks = 0; ks < NK; ++ks) wb[li][ks] = wt[((li * NW + wave) * NK + ks) * 64 + lane];
  for (int e = tid; e < S; e += CF::T) tw[e] = twS[e];
  for (int e = tid; e < 5 * S; e += CF::T) pfl[e] = post[e];

  const long ntiles = (nJ + CF::TJ - 1) / CF::TJ;
  const long t0 = (long)blockIdx.x * tiles_per_block;
  const long t1 = min(ntiles, t0 + tiles_per_block);
  const long N = M * S;
  if (t0 >= t1) return;

  auto dma = [&](float2 *dst, long g0) __attribute__((always_inline)) {
#pragma unroll
    for (int c = wave; c < CF::NDMA; c += NW) {
      long gi = g0 + c * CH + lane * CF::EPL;  // N and g0 are multiples of EPL
      if (gi >= N) gi -= N;
      __builtin_amdgcn_global_load_lds((const void *)(x + gi), (lds_void_t *)(dst + c * CP), 16, 0, 0);
    }
  };

  const int am = n >> 1, apart = n & 1;
  const int v0 = (12 * am + g) * S + wave;

  // K steps [KS0, KS0 + KH) of a tile whose rows are in buffer BUF, accumulated into acc
  auto mfma_half = [&](f32x4 (&acc)[LPW], auto buf_tag, auto half_tag) __attribute__((always_inline)) {
    constexpr int BUF = decltype(buf_tag)::value;
    constexpr int KS0 = decltype(half_tag)::value * KH;
    const float *curf = reinterpret_cast<const float *>(BUF ? raw1 : raw0);
    int vb = v0;
    asm volatile("" : "+v"(vb));
    constexpr bool FAST = (4 * S) % CH == 0;
    const int abase = 2 * ((vb / CH) * CP + vb % CH) + apart;
#pragma unroll
    for (int ks = KS0; ks < KS0 + KH; ++ks) {
#pragma unroll
      for (int li = 0; li < LPW; ++li) {
        float a;
        if constexpr (FAST) {
          a = curf[abase + 2 * (ks * (4 * S / CH) * CP + li * NW)];
        } else {
          const int e = vb + 4 * ks * S + li * NW;
          a = curf[2 * ((e / CH) * CP + e % CH) + apart];
        }
        acc[li] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, wb[li][ks], acc[li], 0, 0, 0);
      }
    }
  };

  // post phase and LPW-point DFTs of a tile's accumulators into the exchange rows
  auto epi_write = [&](const f32x4 (&acc)[LPW]) __attribute__((always_inline)) {
    int tv = tid;
    asm volatile("" : "+v"(tv));
    const int ln = tv & 63, nv = ln & 15, gv = ln >> 4;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      float2 v[LPW];
#pragma unroll
      for (int li = 0; li < LPW; ++li)
        v[li] = cmul(float2{acc[li][2 * h], acc[li][2 * h + 1]}, pfl[(nv % 5) * S + li * NW + wave]);
      dft_r<LPW>(v);
      if (nv < 15) {
        float2 *gr = G + (15 * (2 * gv + h) + nv) * GP + wave;
        gr[0] = v[0];
#pragma unroll
        for (int s1 = 1; s1 < LPW; ++s1) gr[s1 * NW] = cmul(v[s1], tw[(wave * s1) % S]);
      }
    }
  };

  // NW-point DFTs over the waves for each (R, s1) of a tile, stored segment-major
  auto epi_store = [&](long tile, auto part_tag) __attribute__((always_inline)) {
    constexpr bool PART = decltype(part_tag)::value;
    int tv = tid;
    asm volatile("" : "+v"(tv));
    const long R0 = tile * CF::TJ * 5;
#pragma unroll
    for (int c = 0; c < CF::NC; ++c) {
      int e = tv + c * CF::T;
      if constexpr (CF::DUP) {
        if (e >= CF::NTASK) e -= CF::DUPD;
      }
      if (CF::DUP || CF::NTASK % CF::T == 0 || e < CF::NTASK) {
        const int r = e % CF::GR, s1 = e / CF::GR;
        float2 u[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) u[w] = G[r * GP + s1 * NW + w];
        dft_r<NW>(u);
        if (!PART || tile * CF::TJ + r / 5 < nJ) {
#pragma unroll
          for (int s2 = 0; s2 < NW; ++s2)
            store_nt_global(atg + (long)(s1 + LPW * s2) * Mh + R0 + r, u[s2]);
        }
      }
    }
  };

  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;

  // prologue: the first tile's rows, its MFMA chain, the second tile's rows
  dma(raw0, t0 * CF::TJ * 4 * S);
  wait_vmcnt<0>();
  lds_barrier();
  if (t0 + 1 < t1) dma(raw1, (t0 + 1) * CF::TJ * 4 * S);
  f32x4 acc[LPW];
#pragma unroll
  for (int li = 0; li < LPW; ++li) acc[li] = f32x4{0, 0, 0, 0};
  mfma_half(acc, I0{}, I0{});
  mfma_half(acc, I0{}, I1{});
  wait_vmcnt<0>();
  lds_barrier();  // the second tile's rows are in; raw0 is free

  // step t: whole tile t with a tile t + 1 behind it; NB = buffer of tile t + 1.  LATE: this
  // wave runs the epilogue before the MFMA half in the first interval and after it in the second.
  auto step = [&](long tile, auto nb_tag, auto late_tag) __attribute__((always_inline)) {
    constexpr int NB = decltype(nb_tag)::value;
    constexpr bool LATE = decltype(late_tag)::value;
    using BN = std::integral_constant<int, NB>;
    if (tile + 2 < t1) dma(NB ? raw0 : raw1, (tile + 2) * CF::TJ * 4 * S);
    f32x4 accn[LPW];
#pragma unroll
    for (int li = 0; li < LPW; ++li) accn[li] = f32x4{0, 0, 0, 0};
    if constexpr (LATE) {
      epi_write(acc);
      mfma_half(accn, BN{}, I0{});
    } else {
      mfma_half(accn, BN{}, I0{});
      epi_write(acc);
    }
    lds_barrier();  // the exchange rows of tile t are complete
    if constexpr (LATE) {
      mfma_half(accn, BN{}, I1{});
      epi_store(tile, std::false_type{});
    } else {
      epi_store(tile, std::false_type{});
      mfma_half(accn, BN{}, I1{});
    }
    // the loads of tile t + 2 are complete once at most this wave's stores of tile t are
    // outstanding (issued after them); unconditional, so every path back to the loop head waits
    asm volatile("" ::: "memory");
    if constexpr (CF::DUP) wait_vmcnt<CF::NC * NW>();
    else wait_vmcnt<0>();
    lds_barrier();  // tile t + 2's rows are in; the exchange rows and buffer NB are free
#pragma unroll
    for (int li = 0; li < LPW; ++li) acc[li] = accn[li];
  };

  // whole tiles with a successor, two per iteration (static buffer names); then the last tile.
  // One copy of the loop per wave group: a branch inside the loop between the two orders would
  // make the compiler lose track of which buffer a pending load writes (and wait for all of them).
  const long tfull = min(t1, nJ / CF::TJ);
  auto run = [&](auto late_tag) __attribute__((always_inline)) {
    long tile = t0;
    for (; tile + 2 < t1 && tile + 1 < tfull; tile += 2) {
      step(tile, I1{}, late_tag);
      step(tile + 1, I0{}, late_tag);
    }
    if (tile + 1 < t1 && tile < tfull) {
      step(tile, I1{}, late_tag);
      ++tile;
    }
    // tile == t1 - 1: its accumulators are in acc
    epi_write(acc);
    lds_barrier();
    if (tile < tfull) epi_store(tile, std::false_type{});
    else epi_store(tile, std::true_type{});
  };
  if (late) run(std::true_type{});
  else run(std::false_type{});
}

// ---------------------------------------------------------------------------------------------
// Complex128 real-tap filter on MFMA (v_mfma_f64_16x16x4_f64, FP64 products and accumulation).
// The GEMM of each lane i is the one of the FP32 MFMA form with the rows ordered part-major,
//   C[8 part + m][theta + 5D] = sum_k x_i[4 J0 + 12 m + k].part * t_i[theta][k - 4D],
// m < 8, D < 3 (24 row groups J = J0 + 3m + D per tile), K = 80 >= BPAD + 8 in 20 four-deep
// steps.  Lane layout (checked by soik_mfma64_layout_probe): A[row][k] in lane row + 16k,
// B[k][col] in lane col + 16k, D[row][col] in lane col + 16 (row % 4), register row / 4.  With
// part-major rows, lane col + 16g holds m = g (registers 0 and 2: real and imaginary part) and
// m = 4 + g (registers 1 and 3).
//
// Wave w of the NW = S / LPW waves owns the lanes i = NW li + w, li < LPW.  The S-point DFT of
// each output row R then splits as i = NW li + w, s = s1 + LPW s2:
//   X[s] = sum_w w_NW^{w s2} [ w_S^{w s1} sum_li gamma[NW li + w] w_LPW^{li s1} ],
// the LPW-point DFTs in the registers of the wave that holds the lanes, then one exchange
// through LDS and an NW-point DFT per (R, s1), whose NW results go straight to HBM.
//
// A complex128 tile (164 input rows, 82 KB at S = 32) and its output rows (61 KB) do not fit
// twice in LDS, so the input rows live in a ring of RC one-KB chunks (the direct-to-LDS load of
// one wave, each followed by a 16-byte gap): the tile of row groups [24t, 24t + 24) reads chunks
// [AC t, AC t + RC) of the input, slot = chunk mod RC, and the AC chunks that tile t + 1 adds
// overwrite the AC slots that only tile t reads.  Those loads are issued as soon as tile t's
// MFMA block is done and complete under its epilogue.  The B operands (taps, 20 per lane i)
// stay in registers.
// ---------------------------------------------------------------------------------------------
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef double f64x2g __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) f64x2g gf64x2;

// non-temporal store of one complex128 value through a global-address-space pointer: a flat
// store would also count against LDS traffic, and the compiler then waits for the in-flight
// direct-to-LDS loads before it
__device__ __forceinline__ void store_nt_global(gf64x2 *p, double2 v)
{
  __builtin_nontemporal_store(f64x2g{v.x, v.y}, p);
}

template <int S, int LPW> struct Mfma64Cfg {
  static constexpr int TJ = 24;                 // row groups per tile
  static constexpr int NK = 20;                 // four-deep K steps
  static constexpr int K = 4 * NK;
  static constexpr int ROWS = 12 * 7 + K;       // input rows read per tile
  static constexpr int ADV = TJ * 4;            // rows from one tile to the next
  static constexpr int CH = 64;                 // complex128 values per wave-wide load (1 KB)
  static constexpr int CP = CH + 1;             // chunk pitch in LDS: a 16-byte gap
  static constexpr int RC = (ROWS * S + CH - 1) / CH;  // ring chunks: one tile's rows
  static constexpr int AC = ADV * S / CH;       // chunks added per tile
  static constexpr int NWAVES = S / LPW;
  static constexpr int T = NWAVES * 64;
  static constexpr int GR = TJ * 5;             // output rows R per tile
  static constexpr int GP = S + 1;              // exchange row pitch: [R][s1 * NWAVES + w]
  static constexpr int NTASK = GR * LPW;        // (R, s1) tasks of the cross-wave DFT
  static constexpr int NC = (NTASK + T - 1) / T;
  // DUP: the idle lanes of the last task slot repeat tasks of their own wave (every wave then
  // issues the same NC * NWAVES stores per tile); otherwise whole waves would be idle
  static constexpr int DUPD = NC * T - NTASK;
  static constexpr bool DUP = DUPD < 64;
  static_assert(ADV * S % CH == 0 && AC <= RC, "ring shape");
  static_assert(LPW == 2 || LPW == 4 || LPW == 8, "in-register DFT size");
  static constexpr int LDS_BYTES = (RC * CP + GR * GP + S + 5 * S) * 16;
};

constexpr int mfma64_threads(int S, int LPW) { return S / LPW * 64; }

template <int S, int LPW>
__global__ __launch_bounds__(mfma64_threads(S, LPW), mfma64_threads(S, LPW) / 256) void k_filter_mfma64(
    const double2 *__restrict__ x, double2 *__restrict__ at, const double *__restrict__ wt,
    const double2 *__restrict__ post, const double2 *__restrict__ twS, long M, long Mh, long nJ,
    int tiles_per_block, SoiBatch b)
{
  typedef Mfma64Cfg<S, LPW> CF;
  x = batch_ptr(x, b.x);
  at = batch_ptr(at, b.at);
  constexpr int NK = CF::NK, CH = CF::CH, CP = CF::CP, RC = CF::RC, AC = CF::AC;
  constexpr int NW = CF::NWAVES, GP = CF::GP;
  gf64x2 *const atg = (gf64x2 *)at;
  __shared__ __align__(16) double2 ring[RC * CP];
  __shared__ __align__(16) double2 G[CF::GR * GP];
  __shared__ __align__(16) double2 tw[S];
  __shared__ __align__(16) double2 pfl[5 * S];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int n = lane & 15, g = lane >> 4;

  double wb[LPW][NK];  // B operands of this wave's lanes i = NW li + wave
#pragma unroll
  for (int li = 0; li < LPW; ++li)
#pragma unroll
    for (int ks = 0; ks < NK; ++ks) wb[li][ks] = wt[((li * NW + wave) * NK + ks) * 64 + lane];
  for (int e = tid; e < S; e += CF::T) tw[e] = twS[e];
  for (int e = tid; e < 5 * S; e += CF::T) pfl[e] = post[e];

  const long ntiles = (nJ + CF::TJ - 1) / CF::TJ;
  const long t0 = (long)blockIdx.x * tiles_per_block;
  const long t1 = min(ntiles, t0 + tiles_per_block);
  const long N = M * S;
  if (t0 >= t1) return;

  // input chunks [c0, c0 + cnt) into their ring slots (absolute chunk c -> slot c mod RC);
  // a chunk is 64 consecutive values from element 64 c, wrapping at N
  auto dma = [&](long c0, int cnt) {
    int slot = (int)((c0 + wave) % RC);
    for (int c = wave; c < cnt; c += NW) {
      long gi = (c0 + c) * CH + lane;
      if (gi >= N) gi -= N;
      __builtin_amdgcn_global_load_lds((const void *)(x + gi), (lds_void_t *)(ring + slot * CP), 16,
                                       0, 0);
      slot += NW;
      if (slot >= RC) slot -= RC;
    }
  };

  dma(t0 * AC, RC);
  wait_vmcnt<0>();
  lds_barrier();

  // A operand of lane (n, g): row (m, part) = (n & 7, n >> 3), k = g; value index within the
  // tile (12 m + 4 ks + g) * S + i with i = NW li + wave
  const int am = n & 7, apart = n >> 3;
  const int v0 = (12 * am + g) * S + wave;
  int base = (int)((t0 * AC) % RC);  // ring slot of the tile's first chunk
  const double *ringd = reinterpret_cast<const double *>(ring);

  // One tile.  PART: the transform's last tile when it holds fewer than TJ row groups (its
  // stores are predicated per row; it is never followed by another tile of this block).
  // Whole tiles issue the same store instructions in every wave, unconditionally (in DUP
  // mode idle lanes repeat a neighbour's task and store the same value to the same address),
  // so that the compiler's wait tracking sees every one of them and the ring reads of the
  // next tile wait only for the loads issued before them, not for the stores issued after.
  auto body = [&](long tile, auto part_tag) {
    constexpr bool PART = decltype(part_tag)::value;
    const bool more = !PART && tile + 1 < t1;
    // the epilogue's per-thread indices and addresses are recomputed each tile from an opaque
    // copy of the thread index: hoisted out of the tile loop they would stay live through the
    // MFMA block, beside the 20 B operands per lane i
    int tv = tid;
    asm volatile("" : "+v"(tv));
    const int ln = tv & 63, nv = ln & 15, gv = ln >> 4;

    // ---- MFMA: one accumulation chain per lane i, all on v_mfma_f64_16x16x4_f64
    // A operand addresses.  For S >= 16 a K step advances by whole chunks (4S/64) and the lanes
    // i of a wave stay inside one chunk, so a read is the lane's first address of the tile plus
    // a compile-time offset, minus the ring length once the lane's chunk wraps.  (Per-read
    // offsets computed once for the kernel would be loop-invariant and hoisted: 80 live
    // registers.)  Smaller S take the chunk and offset of each read from its value index.
    constexpr int CPB = CP * 16;            // chunk pitch in bytes
    int a0 = 0, a0w = 0, wrap = 0;
    if constexpr (S >= 16) {
      int sb = base + (v0 >> 6);
      if (sb >= RC) sb -= RC;
      a0 = sb * CPB + (v0 & 63) * 16 + apart * 8;
      a0w = a0 - RC * CPB;
      wrap = RC - sb;
    }
    const char *ringb = reinterpret_cast<const char *>(ring);
    auto aread = [&](int ks, int li) -> double {
      if constexpr (S >= 16) {
        constexpr int CS = 4 * S / 64;      // chunks per K step
        const int cs = ks * CS;
        const int ab = cs >= wrap ? a0w : a0;
        return *reinterpret_cast<const double *>(ringb + ab + cs * CPB + li * NW * 16);
      } else {
        const int v = v0 + 4 * ks * S + li * NW;
        int slot = base + (v >> 6);
        if (slot >= RC) slot -= RC;
        return ringd[2 * (slot * CP + (v & 63)) + apart];
      }
    };
    // software pipeline: the A operands of step ks + 1 are read while step ks multiplies; the
    // scheduling barriers keep every read from being hoisted to the top of the block
    f64x4 acc[LPW];
    double acur[LPW];
#pragma unroll
    for (int li = 0; li < LPW; ++li) {
      acc[li] = f64x4{0, 0, 0, 0};
      acur[li] = aread(0, li);
    }
#pragma unroll
    for (int ks = 0; ks < NK; ++ks) {
      double anxt[LPW];
      if (ks + 1 < NK) {
#pragma unroll
        for (int li = 0; li < LPW; ++li) anxt[li] = aread(ks + 1, li);
      }
#pragma unroll
      for (int li = 0; li < LPW; ++li)
        acc[li] = __builtin_amdgcn_mfma_f64_16x16x4f64(acur[li], wb[li][ks], acc[li], 0, 0, 0);
      if (ks + 1 < NK) {
#pragma unroll
        for (int li = 0; li < LPW; ++li) acur[li] = anxt[li];
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    lds_barrier();  // every wave is done reading this tile's rows and the previous tile's G
    if (more) dma((tile + 1) * AC + RC - AC, AC);  // the next tile's new chunks, into dead slots
    base += AC;
    if (base >= RC) base -= RC;

    // ---- post phase and the LPW-point DFTs over this wave's lanes i, twiddled by w_S^{w s1},
    // into the exchange rows: G[R][s1 * NW + w].  Registers h and h + 2 are the real and
    // imaginary part of row m = 4h + g; column nv = theta + 5D (15 is the unused column).
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      double2 v[LPW];
#pragma unroll
      for (int li = 0; li < LPW; ++li)
        v[li] = cmul(double2{acc[li][h], acc[li][h + 2]}, pfl[(nv % 5) * S + li * NW + wave]);
      dft_r<LPW>(v);
      if (nv < 15) {
        double2 *gr = G + (15 * (4 * h + gv) + nv) * GP + wave;
        gr[0] = v[0];
#pragma unroll
        for (int s1 = 1; s1 < LPW; ++s1) gr[s1 * NW] = cmul(v[s1], tw[(wave * s1) % S]);
      }
    }
    lds_barrier();

    // ---- NW-point DFT over the waves for each (R, s1), stored segment-major: X[s1 + LPW s2]
    // of row R = 5 J0 + r goes to alpha_tilde[s1 + LPW s2][R]; consecutive lanes take
    // consecutive r, so each store instruction writes runs of consecutive R
    const long R0 = tile * CF::TJ * 5;
#pragma unroll
    for (int c = 0; c < CF::NC; ++c) {
      int e = tv + c * CF::T;
      if constexpr (CF::DUP) {
        if (e >= CF::NTASK) e -= CF::DUPD;
      }
      if (CF::DUP || CF::NTASK % CF::T == 0 || e < CF::NTASK) {
        const int r = e % CF::GR, s1 = e / CF::GR;
        double2 u[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) u[w] = G[r * GP + s1 * NW + w];
        dft_r<NW>(u);
        if (!PART || tile * CF::TJ + r / 5 < nJ) {
#pragma unroll
          for (int s2 = 0; s2 < NW; ++s2)
            store_nt_global(atg + (long)(s1 + LPW * s2) * Mh + R0 + r, u[s2]);
        }
      }
    }
    // the next tile's chunks are complete once at most this wave's stores of this tile are
    // outstanding (they were issued after the loads).  Unconditional on every path that loops
    // back, so the compiler's wait tracking knows the ring is complete at the next tile.
    if constexpr (!PART) {
      asm volatile("" ::: "memory");
      if constexpr (CF::DUP) wait_vmcnt<CF::NC * NW>();
      else wait_vmcnt<0>();
    }
    lds_barrier();
  };

  // whole tiles in the loop; the transform's partial last tile, if this block has it, after it
  const long tfull = min(t1, nJ / CF::TJ);
  for (long tile = t0; tile < tfull; ++tile) body(tile, std::false_type{});
  if (tfull < t1) body(tfull, std::true_type{});
}

// FP64 B operand table: wt[(i*NK + ks)*64 + lane] = t_i[theta][k - 4D] for k = 4 ks + lane/16
// and col = lane % 16 = theta + 5D (zero for col 15 or a tap index outside [0, BPAD)), from the
// FP64 real-tap table taps[(theta*S + i)*BPAD + b].
__global__ void k_filter_mfma64_table(const double *__restrict__ taps, double *__restrict__ wt, int S,
                                      int BPAD, int NK)
{
  const long total = (long)S * NK * 64;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int lane = (int)(e % 64), ks = (int)(e / 64 % NK), i = (int)(e / (64L * NK));
    const int col = lane % 16, k = 4 * ks + lane / 16;
    const int theta = col % 5, D = col / 5, bb = k - 4 * D;
    const bool on = col < 15 && bb >= 0 && bb < BPAD;
    wt[e] = on ? taps[((long)theta * S + i) * BPAD + bb] : 0.0;
  }
}

template <int S, int LPW> static int launch_mfma64(const SoiFilterArgs *a, hipStream_t st)
{
  typedef Mfma64Cfg<S, LPW> CF;
  static_assert(CF::LDS_BYTES <= 160 * 1024, "complex128 MFMA filter exceeds LDS");
  if ((long)CF::RC * CF::CH > a->M * S) return (int)hipErrorInvalidValue;  // wraps at most once
  const long ntiles = (a->nJ + CF::TJ - 1) / CF::TJ;
  long nb = mfma_blocks_per_member(a);
  if (nb > ntiles) nb = ntiles;
  const int tpb = (int)((ntiles + nb - 1) / nb);
  nb = (ntiles + tpb - 1) / tpb;
  hipLaunchKernelGGL((k_filter_mfma64<S, LPW>), dim3((unsigned)nb, a->b.n > 0 ? a->b.n : 1),
                     dim3(CF::T), 0, st, (const double2 *)a->x, (double2 *)a->at,
                     (const double *)a->wt, (const double2 *)a->post, (const double2 *)a->twS, a->M,
                     a->Mh, a->nJ, tpb, a->b);
  return (int)hipGetLastError();
}

static int filter_mfma64(const SoiFilterArgs *a, hipStream_t st)
{
  switch (a->S) {
    case 4: return launch_mfma64<4, 2>(a, st);
    case 8: return launch_mfma64<8, 2>(a, st);
    case 16: return launch_mfma64<16, 2>(a, st);
    case 32: return launch_mfma64<32, 4>(a, st);
    default: return (int)hipErrorInvalidValue;
  }
}

// FP32 B operand table: wt[(i*NK + ks)*64 + lane] = t_i[theta][k - 4D] for k = 4 ks + lane/16
// and n = lane % 16 = theta + 5D (zero for n = 15 or a tap index outside [0, BPAD)), rounded
// once from the FP64 real-tap table taps[(theta*S + i)*BPAD + b].
__global__ void k_filter_mfma32_table(const double *__restrict__ taps, float *__restrict__ wt, int S,
                                      int BPAD, int NK)
{
  const long total = (long)S * NK * 64;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int lane = (int)(e % 64), ks = (int)(e / 64 % NK), i = (int)(e / (64L * NK));
    const int n = lane % 16, k = 4 * ks + lane / 16;
    const int theta = n % 5, D = n / 5, bb = k - 4 * D;
    const bool on = n < 15 && bb >= 0 && bb < BPAD;
    wt[e] = on ? (float)taps[((long)theta * S + i) * BPAD + bb] : 0.0f;
  }
}

template <int S, int LPW> static int launch_mfma32(const SoiFilterArgs *a, hipStream_t st)
{
  typedef Mfma32Cfg<S, LPW> CF;
  if ((long)CF::NDMA * CF::CH > a->M * S) return (int)hipErrorInvalidValue;  // wraps at most once
  const long ntiles = (a->nJ + CF::TJ - 1) / CF::TJ;
  long nb = 256;
  if (nb > ntiles) nb = ntiles;
  const int tpb = (int)((ntiles + nb - 1) / nb);
  nb = (ntiles + tpb - 1) / tpb;
  static const int pipe = [] {
    const char *e = getenv("SOI_FILTER_MFMA32_PIPE");
    return e ? atoi(e) : 1;
  }();
  auto kern = pipe ? k_filter_mfma32p<S, LPW> : k_filter_mfma32<S, LPW>;
  hipLaunchKernelGGL(kern, dim3((unsigned)nb, a->b.n > 0 ? a->b.n : 1),
                     dim3(CF::T), 0, st, (const float2 *)a->x, (float2 *)a->at,
                     (const float *)a->wt, (const float2 *)a->post, (const float2 *)a->twS, a->M,
                     a->Mh, a->nJ, tpb, a->b);
  return (int)hipGetLastError();
}

static int filter_mfma32(const SoiFilterArgs *a, hipStream_t st)
{
  switch (a->S) {
    case 4: return launch_mfma32<4, 2>(a, st);
    case 8: return launch_mfma32<8, 2>(a, st);
    case 16: return launch_mfma32<16, 2>(a, st);
    case 32: return launch_mfma32<32, 4>(a, st);
    default: return (int)hipErrorInvalidValue;
  }
}

// The MFMA forms apply to the complex64 real-tap (5, 4) filter with S in {4, 8, 16, 32} and at
// most 88 (split-bf16, mode 1) or 72 (FP32, mode 2) padded taps.
extern "C" int soik_filter_mfma_supported(int S, int n_mu, int d_mu, int ct, int BPAD, int mode)
{
  const int maxb = mode == 1 ? 88 : Mfma32Cfg<4, 2>::K - 8;
  return (mode == 1 || mode == 2 || mode == 3) && !ct && n_mu == 5 && d_mu == 4 && BPAD <= maxb &&
         (S == 4 || S == 8 || S == 16 || S == 32);
}

extern "C" size_t soik_filter_mfma_table_bytes(int S, int BPAD, int mode)
{
  (void)BPAD;
  if (mode == 3) return (size_t)S * Mfma64Cfg<4, 2>::NK * 64 * sizeof(double);
  if (mode == 2) return (size_t)S * Mfma32Cfg<4, 2>::NK * 64 * sizeof(float);
  return (size_t)S * MfmaCfg<4>::NCH * 3 * 64 * 16;
}

extern "C" int soik_filter_mfma_table(const double *taps, void *wt, int S, int BPAD, int mode,
                                      hipStream_t st)
{
  if (mode == 3)
    hipLaunchKernelGGL(k_filter_mfma64_table, dim3(64), dim3(256), 0, st, taps, (double *)wt, S,
                       BPAD, Mfma64Cfg<4, 2>::NK);
  else if (mode == 2)
    hipLaunchKernelGGL(k_filter_mfma32_table, dim3(64), dim3(256), 0, st, taps, (float *)wt, S,
                       BPAD, Mfma32Cfg<4, 2>::NK);
  else
    hipLaunchKernelGGL(k_filter_mfma_table, dim3(64), dim3(256), 0, st, taps, (unsigned short *)wt,
                       S, BPAD, MfmaCfg<4>::NCH);
  return (int)hipGetLastError();
}

static int filter_mfma(const SoiFilterArgs *a, hipStream_t st)
{
  switch (a->S) {
    case 4: return launch_mfma<4>(a, st);
    case 8: return launch_mfma<8>(a, st);
    case 16: return launch_mfma<16>(a, st);
    case 32: return launch_mfma<32>(a, st);
    default: return (int)hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------------------------------------
// Lane layout check of v_mfma_f64_16x16x4_f64, the instruction of the complex128 MFMA filter.
// One wave.  Experiments e < 64: A is one-hot at lane e (1.0), B holds lane + 1 in lane, C = 0.
// Experiment 64: small integers in A, B and C (every product and sum exact in FP64).  Output
// out[e*256 + lane*4 + r] = result register r of lane; 65*256 doubles.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_mfma64_layout_probe(double *__restrict__ out)
{
  const int lane = threadIdx.x;
  for (int e = 0; e < 64; ++e) {
    const double a = lane == e ? 1.0 : 0.0;
    const double b = (double)(lane + 1);
    f64x4 acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) out[e * 256 + lane * 4 + r] = acc[r];
  }
  const double a = (double)((lane * 7) % 11 - 5);
  const double b = (double)((lane * 5) % 13 - 6);
  f64x4 acc;
#pragma unroll
  for (int r = 0; r < 4; ++r) acc[r] = (double)(100 * lane + r);
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
#pragma unroll
  for (int r = 0; r < 4; ++r) out[64 * 256 + lane * 4 + r] = acc[r];
}

extern "C" int soik_mfma64_layout_probe(double *out, hipStream_t st)
{
  hipLaunchKernelGGL(k_mfma64_layout_probe, dim3(1), dim3(64), 0, st, out);
  return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Dispatch table
// ---------------------------------------------------------------------------------------------
template <int S, int NMU, int NT, int DMU, int BP, int U, int NP, int W, bool CT, typename C>
static int launch1(const SoiFilterArgs *a, hipStream_t st)
{
  typedef FilterCfg<S, NMU, NT, DMU, BP, U, NP, W, CT, C> CF;
  static_assert(CF::LDS_BYTES <= 160 * 1024, "filter tile exceeds LDS");
  if ((long)CF::XN > a->M * S) return (int)hipErrorInvalidValue;  // tile window wraps at most once
  auto kern = a->b.n > 0 ? k_filter<S, NMU, NT, DMU, BP, U, NP, W, CT, C, true>
                         : k_filter<S, NMU, NT, DMU, BP, U, NP, W, CT, C, false>;
  // tiles of all members (one transform, or the members of a batched launch)
  const long ntiles = (a->nJ + CF::TJ - 1) / CF::TJ * (a->b.n > 0 ? a->b.n : 1);
  // blocks per phase group: enough to fill the CUs, tiles split into contiguous ranges
  long nb = 256L * CF::BPC / CF::NG;
  if (nb < 1) nb = 1;
  if (nb > ntiles) nb = ntiles;
  const int tpb = (int)((ntiles + nb - 1) / nb);
  nb = (ntiles + tpb - 1) / tpb;
  hipLaunchKernelGGL(kern, dim3((unsigned)(nb * CF::NG)), dim3(CF::T), 0,
                     st, (const C *)a->x, (C *)a->at, (const real_t<C> *)a->taps,
                     (const C *)a->post, (const C *)a->twS, a->M, a->Mh, a->nJ, tpb, a->b);
  return (int)hipGetLastError();
}

// FP64 (the drop-in) or FP32 (the precision form: complex64 data, taps and twiddles rounded from
// their FP64 values)
template <int S, int NMU, int NT, int DMU, int BP, int U, int NP, int W, bool CT>
static int launch(const SoiFilterArgs *a, hipStream_t st)
{
  if (a->f32) return launch1<S, NMU, NT, DMU, BP, U, NP, W, CT, float2>(a, st);
  return launch1<S, NMU, NT, DMU, BP, U, NP, W, CT, double2>(a, st);
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
  const int es = a->f32 ? 8 : 16, ch = 64 * (16 / es);
  const long xn = ((long)rows * a->S + ch - 1) / ch * ch;
  const long gn = (long)TJ * a->NT * (a->S + 1);
  const long buf = xn > gn ? xn : gn;
  const int db = (2 * buf + a->S) * es <= 160 * 1024;
  return (size_t)((db ? 2 : 1) * buf + a->S) * es;
}

#define RT_CASE(SV, BPV, WV)                                                        \
  if (a->S == SV && a->BP == BPV && a->W == WV) return launch<SV, 5, 5, 4, BPV, 8, 2, WV, false>(a, st);
#define CT54_CASE(SV, NTV, BPV, WV)                                                 \
  if (a->S == SV && a->BP == BPV) return launch<SV, 5, NTV, 4, BPV, 8, 4, WV, true>(a, st);
#define CT87_CASE(SV, NTV, BPV, WV)                                                 \
  if (a->S == SV && a->BP == BPV) return launch<SV, 8, NTV, 7, BPV, 8, 4, WV, true>(a, st);

extern "C" int soik_filter_sdft(const SoiFilterArgs *a, hipStream_t st)
{
  if (a->wt && a->mfma == 3 && !a->f32) return filter_mfma64(a, st);
  if (a->f32 && a->wt) return a->mfma == 2 ? filter_mfma32(a, st) : filter_mfma(a, st);
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
