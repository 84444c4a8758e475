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
// Layout and twiddles:
//   * element x lives at LDS index pad(x) = x + x/16: one spare element every
//     16 makes the stride-RAD writes of the first pass and all later passes
//     free of bank conflicts for 8- and 16-byte elements;
//   * W_N^x = T1[x >> S] * T2[x & (2^S-1)] from two small LDS tables
//     (sqrt(N) entries each, filled once per workgroup from the correctly
//     rounded global table), so no pass waits on global memory;
//   * radix = largest power of two <= 32 dividing the per-thread element
//     count E = P/T (radix-32 first: 8192 points with E = 32 take 3 passes).
// INV selects exp(+...) (unnormalised).
#pragma once
#include "rks_math.h"

namespace rksfft {

__device__ __forceinline__ constexpr int pad(int x) { return x + (x >> 4); }
// LDS elements needed for a padded P-point buffer
__host__ __device__ constexpr int padded_size(int P) { return P + (P >> 4) + 1; }

// two-level twiddle table geometry for an N-point base table
template <int LOG2N>
struct TwSplit {
  static constexpr int S = (LOG2N + 1) / 2;
  static constexpr int N1 = (1 << LOG2N) >> S;  // T1[i] = W_N^(i << S)
  static constexpr int N2 = 1 << S;              // T2[j] = W_N^j
  static constexpr int SIZE = N1 + N2;
};

// fill the LDS tables from the global table tw[x] = exp(-2 pi i x / N)
template <typename R, int LOG2N, int T>
__device__ __forceinline__ void load_twiddles(cx<R>* twl, const cx<R>* __restrict__ tw) {
  using TS = TwSplit<LOG2N>;
  for (int i = threadIdx.x; i < TS::SIZE; i += T)
    twl[i] = (i < TS::N1) ? tw[i << TS::S] : tw[i - TS::N1];
}

// W_N^x, 0 <= x < N
template <typename R, int LOG2N>
__device__ __forceinline__ cx<R> twiddle(const cx<R>* twl, int x) {
  using TS = TwSplit<LOG2N>;
  const cx<R> a = twl[x >> TS::S];
  const cx<R> b = twl[TS::N1 + (x & (TS::N2 - 1))];
  return a * b;
}

// exp(-2 pi i k / 32), k = 0..15, correctly rounded
__device__ __forceinline__ double c32(int k) {
  const double v[16] = {1.0, 0.9807852804032304491262, 0.9238795325112867561282, 0.8314696123025452370788,
                        0.7071067811865475244008, 0.5555702330196022247428, 0.3826834323650897717285,
                        0.1950903220161282678483, 0.0, -0.1950903220161282678483, -0.3826834323650897717285,
                        -0.5555702330196022247428, -0.7071067811865475244008, -0.8314696123025452370788,
                        -0.9238795325112867561282, -0.9807852804032304491262};
  return v[k];
}
__device__ __forceinline__ double s32(int k) {
  const double v[16] = {0.0, 0.1950903220161282678483, 0.3826834323650897717285, 0.5555702330196022247428,
                        0.7071067811865475244008, 0.8314696123025452370788, 0.9238795325112867561282,
                        0.9807852804032304491262, 1.0, 0.9807852804032304491262, 0.9238795325112867561282,
                        0.8314696123025452370788, 0.7071067811865475244008, 0.5555702330196022247428,
                        0.3826834323650897717285, 0.1950903220161282678483};
  return v[k];
}

// d * W_len^i (forward: exp(-2 pi i i/len); inverse: conjugate)
template <int LEN, bool INV, typename R>
__device__ __forceinline__ cx<R> tw_const(int i, cx<R> d) {
  const int k = i * (32 / LEN);
  if (k == 0) return d;
  if (k == 8) return mul_mi<INV>(d);
  const R c = (R)c32(k);
  const R s = INV ? (R)s32(k) : (R)(-s32(k));
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
        if (len == 32) v[blk + i + half] = tw_const<32, INV>(i, d);
        else if (len == 16) v[blk + i + half] = tw_const<16, INV>(i, d);
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

template <int E, int MAXR>
struct pick_radix {
  static constexpr int v = (E % 32 == 0 && MAXR >= 32) ? 32
                           : (E % 16 == 0 && MAXR >= 16) ? 16
                           : (E % 8 == 0 && MAXR >= 8) ? 8
                           : (E % 4 == 0 && MAXR >= 4) ? 4 : 2;
};

// v[r] *= W_N^(r step) (conjugate for INV), r = 1..RAD-1.  FP64: powers of
// the one table value w = W_N^step by squaring and products p[r] = p[hb]
// p[r - hb] (hb the highest power of two <= r), at most four multiplications
// deep, so a butterfly reads two table entries instead of 2 (RAD - 1).  FP32
// keeps one table lookup per twiddle (the products' rounding would show).
template <int RAD, bool INV, int LOG2N, typename R>
__device__ __forceinline__ void apply_twiddles(cx<R> (&v)[RAD], const cx<R>* twl, int step) {
  if constexpr (sizeof(R) == sizeof(double)) {
    cx<R> p[RAD];
    p[1] = twiddle<R, LOG2N>(twl, step);
#pragma unroll
    for (int r = 2; r < RAD; ++r) {
      const int hb = (r >= 8) ? 8 : (r >= 4) ? 4 : (r >= 2) ? 2 : 1;
      p[r] = (r == hb) ? p[r / 2] * p[r / 2] : p[hb] * p[r - hb];
    }
#pragma unroll
    for (int r = 1; r < RAD; ++r) v[r] = cmul_tw<INV>(v[r], p[r]);
  } else {
#pragma unroll
    for (int r = 1; r < RAD; ++r) v[r] = cmul_tw<INV>(v[r], twiddle<R, LOG2N>(twl, r * step));
  }
}

// One Stockham pass.  Twiddle W_{NS*RAD}^{r k} = W_N^{r k (P/(NS RAD)) TSCALE}, where
// TSCALE = N / P maps a P-point transform onto the N-point table.
// t: the thread's index among the T threads working on this transform.
template <typename R, int P, int T, int NS, int RAD, bool INV, int LOG2N, int TSCALE>
__device__ __forceinline__ void pass(cx<R>* buf, const cx<R>* twl, int t) {
  constexpr int NB = P / (RAD * T);
  static_assert(NB >= 1, "too few elements per thread");
  cx<R> v[NB][RAD];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * T;
#pragma unroll
    for (int r = 0; r < RAD; ++r) v[b][r] = buf[pad(j + r * (P / RAD))];
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * T;
    const int k = j & (NS - 1);
    if (NS > 1) apply_twiddles<RAD, INV, LOG2N>(v[b], twl, k * (P / (NS * RAD)) * TSCALE);
    dft_small<RAD, INV>(v[b]);
    const int base = (j - k) * RAD + k;
#pragma unroll
    for (int r = 0; r < RAD; ++r) buf[pad(base + r * NS)] = v[b][r];
  }
  __syncthreads();
}

// The last pass of a transform with its outputs handed to post(x, value)
// (x the natural-order index) instead of written to LDS.
template <typename R, int P, int T, int NS, int RAD, bool INV, int LOG2N, int TSCALE, typename Post>
__device__ __forceinline__ void last_pass(const cx<R>* buf, const cx<R>* twl, int t, Post post) {
  constexpr int NB = P / (RAD * T);
  static_assert(NB >= 1, "too few elements per thread");
  static_assert(NS * RAD == P, "not the last pass");
  cx<R> v[NB][RAD];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * T;
#pragma unroll
    for (int r = 0; r < RAD; ++r) v[b][r] = buf[pad(j + r * (P / RAD))];
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * T;  // j < NS: the butterfly's outputs land at j + r*NS
    if (NS > 1) apply_twiddles<RAD, INV, LOG2N>(v[b], twl, j * (P / (NS * RAD)) * TSCALE);
    dft_small<RAD, INV>(v[b]);
#pragma unroll
    for (int r = 0; r < RAD; ++r) post(j + r * NS, v[b][r]);
  }
}

template <typename R, int P, int T, int NS, bool INV, int LOG2N, int TSCALE, int MAXR, typename Post>
__device__ __forceinline__ void chain_post(cx<R>* buf, const cx<R>* twl, int t, Post post) {
  constexpr int E = P / T;
  constexpr int R0 = pick_radix<E, MAXR>::v;
  constexpr int REM = P / NS;
  constexpr int RAD = (R0 <= REM) ? R0 : REM;
  if constexpr (NS * RAD == P) {
    last_pass<R, P, T, NS, RAD, INV, LOG2N, TSCALE>(buf, twl, t, post);
  } else {
    pass<R, P, T, NS, RAD, INV, LOG2N, TSCALE>(buf, twl, t);
    chain_post<R, P, T, NS * RAD, INV, LOG2N, TSCALE, MAXR>(buf, twl, t, post);
  }
}

template <typename R, int P, int T, int NS, bool INV, int LOG2N, int TSCALE, int MAXR>
__device__ __forceinline__ void chain(cx<R>* buf, const cx<R>* twl, int t) {
  if constexpr (NS < P) {
    constexpr int E = P / T;
    constexpr int R0 = pick_radix<E, MAXR>::v;
    constexpr int REM = P / NS;
    constexpr int RAD = (R0 <= REM) ? R0 : REM;
    pass<R, P, T, NS, RAD, INV, LOG2N, TSCALE>(buf, twl, t);
    chain<R, P, T, NS * RAD, INV, LOG2N, TSCALE, MAXR>(buf, twl, t);
  }
}

// Full P-point transform in LDS (padded layout).  twl: LDS tables of the
// 2^LOG2N-point base table (P divides 2^LOG2N).  Callers synchronise before.
template <typename R, int P, int T, bool INV, int LOG2N, int MAXR = 32>
__device__ __forceinline__ void fft(cx<R>* buf, const cx<R>* twl) {
  static_assert((P & (P - 1)) == 0, "power of two");
  static_assert(P / T >= 2, "need at least two elements per thread");
  chain<R, P, T, 1, INV, LOG2N, (1 << LOG2N) / P, MAXR>(buf, twl, threadIdx.x);
}

// Full P-point transform whose last pass hands each output (natural-order
// index x, value) to post instead of storing it: post may write LDS (only the
// index it is given) or global memory.  Ends after the last pass's reads are
// synchronised; callers synchronise before reading what post wrote to LDS.
template <typename R, int P, int T, bool INV, int LOG2N, int MAXR, typename Post>
__device__ __forceinline__ void fft_post(cx<R>* buf, const cx<R>* twl, Post post) {
  static_assert((P & (P - 1)) == 0, "power of two");
  static_assert(P / T >= 2, "need at least two elements per thread");
  chain_post<R, P, T, 1, INV, LOG2N, (1 << LOG2N) / P, MAXR>(buf, twl, threadIdx.x, post);
}

// G independent P-point transforms by one workgroup of G*TG threads: threads
// [g*TG, (g+1)*TG) transform the padded buffer buf + g*padded_size(P).  Every
// pass synchronises the whole workgroup, so all threads must call this.
template <typename R, int P, int TG, int G, bool INV, int LOG2N, int MAXR = 32>
__device__ __forceinline__ void fft_multi(cx<R>* buf, const cx<R>* twl) {
  static_assert((P & (P - 1)) == 0, "power of two");
  static_assert(P / TG >= 2, "need at least two elements per thread");
  const int g = threadIdx.x / TG;
  chain<R, P, TG, 1, INV, LOG2N, (1 << LOG2N) / P, MAXR>(buf + g * padded_size(P), twl, threadIdx.x % TG);
}

}  // namespace rksfft
