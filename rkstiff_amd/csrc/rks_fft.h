// LDS-resident Stockham FFT for one workgroup.
//
// The transform of P = 2^k complex points lives in LDS (natural order in and
// out).  Each pass (NS, RAD) is done by all T threads: a thread loads the
// RAD*NB elements of its NB butterflies into registers, the workgroup
// synchronises, then each butterfly is twiddled, transformed in registers and
// written back in Stockham autosort order:
//   read  buf[j + r*P/RAD]                          (r < RAD, j < P/RAD)
//   twiddle by W_{NS*RAD}^{r*(j mod NS)}
//   write buf[(j/NS)*NS*RAD + (j mod NS) + r*NS]
// Radices are the largest power of two <= 16 that divides the per-thread
// element count E = P/T, chosen at compile time.  Twiddles come from a table
// tw[x] = exp(-2 pi i x / TN) of the largest transform size TN in use; a
// P-point pass reads it with stride TN/P.  INV selects exp(+...) (unnormalised).
#pragma once
#include "rks_math.h"

namespace rksfft {

// exp(-2 pi i k / 16), k = 0..7, correctly rounded
__device__ __forceinline__ double c16(int k) {
  const double v[8] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                       0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613};
  return v[k];
}
__device__ __forceinline__ double s16(int k) {
  const double v[8] = {0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
                       1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173};
  return v[k];
}

// d * W_len^i (forward: exp(-2 pi i i/len); inverse: conjugate)
template <int LEN, bool INV, typename R>
__device__ __forceinline__ cx<R> tw_const(int i, cx<R> d) {
  const int k = i * (16 / LEN);
  if (k == 0) return d;
  if (k == 4) return mul_mi<INV>(d);
  const R c = (R)c16(k);
  const R s = INV ? (R)s16(k) : (R)(-s16(k));
  return mkc(d.re * c - d.im * s, d.re * s + d.im * c);
}

template <int RAD>
__device__ constexpr int bitrev(int i) {
  int r = 0;
  for (int b = 1; b < RAD; b <<= 1) {
    r = (r << 1) | (i & 1);
    i >>= 1;
  }
  return r;
}

// In-register DFT of RAD points (radix-2 decimation in frequency).
template <int RAD, bool INV, typename R>
__device__ __forceinline__ void dft_small(cx<R> (&v)[RAD]) {
#pragma unroll
  for (int len = RAD; len >= 2; len >>= 1) {
    const int half = len / 2;
#pragma unroll
    for (int blk = 0; blk < RAD; blk += len) {
#pragma unroll
      for (int i = 0; i < half; ++i) {
        cx<R> a = v[blk + i], b = v[blk + i + half];
        v[blk + i] = a + b;
        cx<R> d = a - b;
        if (len == 16) v[blk + i + half] = tw_const<16, INV>(i, d);
        else if (len == 8) v[blk + i + half] = tw_const<8, INV>(i, d);
        else if (len == 4) v[blk + i + half] = tw_const<4, INV>(i, d);
        else v[blk + i + half] = d;
      }
    }
  }
  cx<R> t[RAD];
#pragma unroll
  for (int i = 0; i < RAD; ++i) t[i] = v[bitrev<RAD>(i)];
#pragma unroll
  for (int i = 0; i < RAD; ++i) v[i] = t[i];
}

template <int E>
struct pick_radix {
  static constexpr int v = (E % 16 == 0) ? 16 : (E % 8 == 0) ? 8 : (E % 4 == 0) ? 4 : 2;
};

template <typename R, int P, int T, int NS, int RAD, bool INV>
__device__ __forceinline__ void pass(cx<R>* buf, const cx<R>* __restrict__ tw, int tws) {
  constexpr int NB = P / (RAD * T);
  static_assert(NB >= 1, "too few elements per thread");
  const int t = threadIdx.x;
  cx<R> v[NB][RAD];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * T;
#pragma unroll
    for (int r = 0; r < RAD; ++r) v[b][r] = buf[j + r * (P / RAD)];
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * T;
    const int k = j & (NS - 1);
    if (NS > 1) {
      const int step = k * (P / (NS * RAD)) * tws;
#pragma unroll
      for (int r = 1; r < RAD; ++r) v[b][r] = cmul_tw<INV>(v[b][r], tw[r * step]);
    }
    dft_small<RAD, INV>(v[b]);
    const int base = (j - k) * RAD + k;
#pragma unroll
    for (int r = 0; r < RAD; ++r) buf[base + r * NS] = v[b][r];
  }
  __syncthreads();
}

template <typename R, int P, int T, int NS, bool INV>
__device__ __forceinline__ void chain(cx<R>* buf, const cx<R>* __restrict__ tw, int tws) {
  if constexpr (NS < P) {
    constexpr int E = P / T;
    constexpr int R0 = pick_radix<E>::v;
    constexpr int REM = P / NS;
    constexpr int RAD = (R0 <= REM) ? R0 : REM;
    pass<R, P, T, NS, RAD, INV>(buf, tw, tws);
    chain<R, P, T, NS * RAD, INV>(buf, tw, tws);
  }
}

// Full P-point transform in LDS.  Entry and exit are workgroup-synchronised
// by the caller/pass barriers.
template <typename R, int P, int T, bool INV>
__device__ __forceinline__ void fft(cx<R>* buf, const cx<R>* __restrict__ tw, int tws) {
  static_assert((P & (P - 1)) == 0, "power of two");
  static_assert(P / T >= 2, "need at least two elements per thread");
  chain<R, P, T, 1, INV>(buf, tw, tws);
}

}  // namespace rksfft
