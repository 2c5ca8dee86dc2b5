// soi_common.hpp -- device helpers shared by the SOI kernels (gfx950, wave64).
//
// Complex arithmetic on HIP vector types, fixed-size forward DFTs (sign -1, unscaled: the
// convention of the reference's DFTI descriptors, pfft.c:224-228) and an in-place LDS Stockham
// transform used for the S-point row DFTs and the M_hat-point segment FFT passes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "soi_kernels.h"

namespace soi {

template <typename T> struct cplx;
template <> struct cplx<double> { typedef double2 t; };
template <> struct cplx<float> { typedef float2 t; };
template <typename C> struct real_of;
template <> struct real_of<double2> { typedef double t; };
template <> struct real_of<float2> { typedef float t; };
template <typename C> using real_t = typename real_of<C>::t;

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b)
{
  return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// W_inv[m] * y in FP64, rounded to the precision of y (the demodulation gain amplifies error
// near segment edges, so the product is formed in FP64 in both forms)
template <typename C> __device__ __forceinline__ C demod(double2 w, C y)
{
  const double yx = (double)y.x, yy = (double)y.y;
  return C{(real_t<C>)(w.x * yx - w.y * yy), (real_t<C>)(w.x * yy + w.y * yx)};
}
// Streaming store of one complex value (one 8- or 16-byte non-temporal vector store): for
// outputs that are not read again before they leave the caches
template <typename C> __device__ __forceinline__ void store_nt(C *p, C v)
{
  typedef real_t<C> R;
  typedef R v2 __attribute__((ext_vector_type(2)));
  v2 t = {v.x, v.y};
  __builtin_nontemporal_store(t, reinterpret_cast<v2 *>(p));
}
// multiply by -i
template <typename C> __device__ __forceinline__ C cmul_mi(C a) { return C{a.y, -a.x}; }
template <typename C, typename S> __device__ __forceinline__ C cscale(C a, S s) { return C{a.x * s, a.y * s}; }

// ---------------------------------------------------------------------------------------------
// Forward DFTs of size R in registers.
// ---------------------------------------------------------------------------------------------
template <typename C> __device__ __forceinline__ void dft2(C &a0, C &a1)
{
  C t = a0;
  a0 = cadd(t, a1);
  a1 = csub(t, a1);
}

template <typename C> __device__ __forceinline__ void dft4(C &a0, C &a1, C &a2, C &a3)
{
  C t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = csub(a1, a3);
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  C mt3 = cmul_mi(t3);
  a1 = cadd(t1, mt3);
  a3 = csub(t1, mt3);
}

template <typename C> __device__ __forceinline__ void dft3(C &a0, C &a1, C &a2)
{
  typedef decltype(a0.x) T;
  const T h = (T)0.86602540378443864676372317075294;  // sqrt(3)/2
  C t = cadd(a1, a2);
  C d = csub(a1, a2);
  C m = C{a0.x - (T)0.5 * t.x, a0.y - (T)0.5 * t.y};
  a0 = cadd(a0, t);
  // X1 = m - i*h*d, X2 = m + i*h*d
  C u = C{h * d.y, -h * d.x};
  a1 = cadd(m, u);
  a2 = csub(m, u);
}

template <typename C> __device__ __forceinline__ void dft5(C &a0, C &a1, C &a2, C &a3, C &a4)
{
  typedef decltype(a0.x) T;
  const T c1 = (T)0.30901699437494742410229341718282;   // cos(2pi/5)
  const T c2 = (T)-0.80901699437494742410229341718282;  // cos(4pi/5)
  const T s1 = (T)0.95105651629515357211643933337938;   // sin(2pi/5)
  const T s2 = (T)0.58778525229247312916870595463907;   // sin(4pi/5)
  C b1 = cadd(a1, a4), b2 = cadd(a2, a3), d1 = csub(a1, a4), d2 = csub(a2, a3);
  C x0 = a0;
  C p1 = C{x0.x + c1 * b1.x + c2 * b2.x, x0.y + c1 * b1.y + c2 * b2.y};
  C p2 = C{x0.x + c2 * b1.x + c1 * b2.x, x0.y + c2 * b1.y + c1 * b2.y};
  C q1 = C{s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y};
  C q2 = C{s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y};
  a0 = C{x0.x + b1.x + b2.x, x0.y + b1.y + b2.y};
  // X1 = p1 - i q1, X4 = p1 + i q1, X2 = p2 - i q2, X3 = p2 + i q2
  a1 = C{p1.x + q1.y, p1.y - q1.x};
  a4 = C{p1.x - q1.y, p1.y + q1.x};
  a2 = C{p2.x + q2.y, p2.y - q2.x};
  a3 = C{p2.x - q2.y, p2.y + q2.x};
}

template <typename C> __device__ __forceinline__ void dft8(C *a)
{
  typedef decltype(a[0].x) T;
  const T r = (T)0.70710678118654752440084436210485;
  dft4(a[0], a[2], a[4], a[6]);
  dft4(a[1], a[3], a[5], a[7]);
  // twiddles w8^k on odd part: w8 = (1-i)/sqrt2, w8^2 = -i, w8^3 = -(1+i)/sqrt2
  C o1 = C{r * (a[3].x + a[3].y), r * (a[3].y - a[3].x)};
  C o2 = cmul_mi(a[5]);
  C o3 = C{r * (a[7].y - a[7].x), -r * (a[7].x + a[7].y)};
  C e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  a[0] = cadd(e0, a[1]);
  a[4] = csub(e0, a[1]);
  a[1] = cadd(e1, o1);
  a[5] = csub(e1, o1);
  a[2] = cadd(e2, o2);
  a[6] = csub(e2, o2);
  a[3] = cadd(e3, o3);
  a[7] = csub(e3, o3);
}

template <int R, typename C> __device__ __forceinline__ void dft_r(C *a)
{
  if constexpr (R == 2) dft2(a[0], a[1]);
  else if constexpr (R == 3) dft3(a[0], a[1], a[2]);
  else if constexpr (R == 4) dft4(a[0], a[1], a[2], a[3]);
  else if constexpr (R == 5) dft5(a[0], a[1], a[2], a[3], a[4]);
  else if constexpr (R == 8) dft8(a);
}


// ---------------------------------------------------------------------------------------------
// In-place LDS Stockham autosort transform of `count` sequences of length L.
//   element j of sequence q lives at buf[q*qs + j*es]
//   tw[m] = e^{-2 pi i m / L}, m < L (global memory, cached)
// Pass with radix R at running span Ns: butterfly j reads j + k*L/R, multiplies input k by
// w_{Ns R}^{k (j mod Ns)} and writes (j/Ns)*Ns*R + (j mod Ns) + k*Ns.  Every thread stages its
// butterflies in registers between two barriers, so at most VMAX values per thread per pass.
// ---------------------------------------------------------------------------------------------
template <int R, int VMAX, typename C>
__device__ __forceinline__ void stockham_pass(C *buf, int L, int count, int qs, int es, int Ns,
                                              const C *__restrict__ tw)
{
  constexpr int NB = VMAX / R;
  const int nbf = L / R;
  const int total = nbf * count;
  const int tid = threadIdx.x, nt = blockDim.x;
  C v[NB][R];
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    int g = tid + c * nt;
    if (g < total) {
      int q = g / nbf, j = g - q * nbf;
      const C *src = buf + q * qs + j * es;
#pragma unroll
      for (int k = 0; k < R; ++k) v[c][k] = src[k * nbf * es];
    }
  }
  __syncthreads();
  const int step = L / (Ns * R);
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    int g = tid + c * nt;
    if (g < total) {
      int q = g / nbf, j = g - q * nbf;
      int jm = j % Ns;
      if (Ns > 1) {
#pragma unroll
        for (int k = 1; k < R; ++k) v[c][k] = cmul(v[c][k], tw[k * jm * step]);
      }
      dft_r<R>(v[c]);
      C *dst = buf + q * qs + ((j - jm) * R + jm) * es;
#pragma unroll
      for (int k = 0; k < R; ++k) dst[k * Ns * es] = v[c][k];
    }
  }
  __syncthreads();
}

// Radix schedule of a length L = 2^a 3^b 5^c: radix-8 passes, then 4 or 2 for the rest of the
// power of two, then 5s, then 3s.  Computed from L in uniform (scalar) arithmetic.
__device__ __forceinline__ int radix_at(int L, int pass)
{
  int p2 = 0, m = L;
  while ((m & 1) == 0) { m >>= 1; ++p2; }
  int n8 = p2 / 3, rem2 = p2 - 3 * n8;
  int n5 = 0;
  while (m % 5 == 0) { m /= 5; ++n5; }
  int n3 = 0;
  while (m % 3 == 0) { m /= 3; ++n3; }
  if (pass < n8) return 8;
  pass -= n8;
  if (rem2) {
    if (pass == 0) return rem2 == 2 ? 4 : 2;
    pass -= 1;
  }
  if (pass < n5) return 5;
  pass -= n5;
  if (pass < n3) return 3;
  return 0;
}

template <int VMAX, typename C>
__device__ __forceinline__ void lds_fft(C *buf, int L, int count, int qs, int es, const C *__restrict__ tw)
{
  int Ns = 1;
  for (int p = 0;; ++p) {
    const int R = radix_at(L, p);
    if (R == 0) break;
    switch (R) {
      case 8: stockham_pass<8, VMAX>(buf, L, count, qs, es, Ns, tw); break;
      case 5: stockham_pass<5, VMAX>(buf, L, count, qs, es, Ns, tw); break;
      case 4: stockham_pass<4, VMAX>(buf, L, count, qs, es, Ns, tw); break;
      case 3: stockham_pass<3, VMAX>(buf, L, count, qs, es, Ns, tw); break;
      default: stockham_pass<2, VMAX>(buf, L, count, qs, es, Ns, tw); break;
    }
    Ns *= R;
  }
}

// Buffer of this block's transform in a batched launch (blockIdx.y indexes the table), or the
// directly passed pointer for a single transform.
template <typename T> __device__ __forceinline__ T *batch_ptr(T *p, void *const *tab)
{
  return tab ? reinterpret_cast<T *>(tab[blockIdx.y]) : p;
}

// Barrier for LDS hand-off only: waits for this wave's LDS operations, not for its global loads
// and stores, which stay in flight across it.
__device__ __forceinline__ void lds_barrier()
{
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// XCD-aware block index: hardware dispatches consecutive workgroups round-robin over the 8 XCDs.
// Remapping makes logically adjacent blocks (which share table rows and cache lines) land on the
// same XCD and its L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks)
{
  const int NX = 8;
  if (nblocks % NX) return bid;
  return (bid % NX) * (nblocks / NX) + bid / NX;
}

}  // namespace soi
