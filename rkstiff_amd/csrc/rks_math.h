// Complex arithmetic for the stepping kernels.
//
// Two flavours:
//  * cx<R> operators: plain complex arithmetic for stage combinations and FFTs
//    (the compiler may contract to FMA; rounding is not part of any contract).
//  * np_* functions (FP64): the rounding conventions of NumPy's complex ufuncs
//    on x86-64 (NumPy 2.x), used by the psi-coefficient generator so that the
//    cancellation-amplified coefficients near modecutoff follow the reference:
//      multiply  re = fma(ar, br, -(ai*bi)), im = fma(ar, bi, ai*br)
//      divide    Smith's algorithm
//      z**3      (z*z)*z without fused multiply-add
//      exp       exp(re) * (cos(im), sin(im))
#pragma once
#include <hip/hip_runtime.h>

template <typename R>
struct __attribute__((aligned(2 * sizeof(R)))) cx {
  R re, im;
};

template <typename R>
__device__ __forceinline__ cx<R> mkc(R a, R b) {
  cx<R> c;
  c.re = a;
  c.im = b;
  return c;
}
template <typename R>
__device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return mkc(a.re + b.re, a.im + b.im); }
template <typename R>
__device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return mkc(a.re - b.re, a.im - b.im); }
template <typename R>
__device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b) {
  return mkc(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// FP64: the fused form NumPy's complex multiply uses (re = fma(ar, br, -(ai bi)), im = fma(ar, bi, ai br))
template <>
__device__ __forceinline__ cx<double> operator*(cx<double> a, cx<double> b) {
  return mkc(__builtin_fma(a.re, b.re, -(a.im * b.im)), __builtin_fma(a.re, b.im, a.im * b.re));
}
template <typename R>
__device__ __forceinline__ cx<R> operator*(R s, cx<R> a) { return mkc(s * a.re, s * a.im); }
template <typename R>
__device__ __forceinline__ cx<R> conjc(cx<R> a) { return mkc(a.re, -a.im); }
// a * conj(b) if inverse else a * b
template <bool CONJ, typename R>
__device__ __forceinline__ cx<R> cmul_tw(cx<R> a, cx<R> w) {
  if (CONJ) return mkc(a.re * w.re + a.im * w.im, a.im * w.re - a.re * w.im);
  return a * w;
}
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename R>
__device__ __forceinline__ cx<R> mul_mi(cx<R> a) {
  if (INV) return mkc(-a.im, a.re);
  return mkc(a.im, -a.re);
}

template <typename A, typename B>
__device__ __forceinline__ cx<A> cvt(cx<B> v) { return mkc((A)v.re, (A)v.im); }

// ---------------------------------------------------------------------------
// NumPy-convention FP64 complex arithmetic (coefficient generation only)
// ---------------------------------------------------------------------------
typedef cx<double> z64;

__device__ __forceinline__ z64 np_mul(z64 a, z64 b) {
#pragma clang fp contract(off)
  return mkc(fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re));
}
__device__ __forceinline__ z64 np_mul_nofma(z64 a, z64 b) {
#pragma clang fp contract(off)
  double rr = a.re * b.re;
  double ii = a.im * b.im;
  double ri = a.re * b.im;
  double ir = a.im * b.re;
  return mkc(rr - ii, ri + ir);
}
__device__ __forceinline__ z64 np_div(z64 a, z64 b) {
#pragma clang fp contract(off)
  double abr = fabs(b.re), abi = fabs(b.im);
  if (abr >= abi) {
    if (abr == 0.0 && abi == 0.0) return mkc(a.re / abr, a.im / abr);
    double rat = b.im / b.re;
    double scl = 1.0 / (b.re + b.im * rat);
    return mkc((a.re + a.im * rat) * scl, (a.im - a.re * rat) * scl);
  }
  double rat = b.re / b.im;
  double scl = 1.0 / (b.im + b.re * rat);
  return mkc((a.re * rat + a.im) * scl, (a.im * rat - a.re) * scl);
}
// real scalar s times complex a (NumPy promotes s to s+0j; the FMA form rounds once per part)
__device__ __forceinline__ z64 np_rmul(double s, z64 a) {
#pragma clang fp contract(off)
  return mkc(s * a.re, s * a.im);
}
// complex divided by a real scalar d: Smith with b.im = 0, i.e. multiply by 1/d
__device__ __forceinline__ z64 np_rdiv(z64 a, double d) {
#pragma clang fp contract(off)
  double scl = 1.0 / d;
  return mkc(a.re * scl, a.im * scl);
}
__device__ __forceinline__ z64 np_exp(z64 z) {
  double e = exp(z.re);
  if (z.im == 0.0) return mkc(e, z.im);
  double s, c;
  sincos(z.im, &s, &c);
  return mkc(e * c, e * s);
}

// psi_r(w) from the closed forms of etd.py:134-182 in NumPy operation order:
//   psi1 = (exp(w) - 1) / w
//   psi2 = 2 * (exp(w) - 1 - w) / w**2
//   psi3 = 6 * (exp(w) - 1 - w - w**2 / 2) / w**3
__device__ __forceinline__ void np_psi123_e(z64 w, z64 e, int need3, z64* p1, z64* p2, z64* p3) {
#pragma clang fp contract(off)
  z64 em1 = mkc(e.re - 1.0, e.im);
  *p1 = np_div(em1, w);
  z64 w2 = np_mul(w, w);
  z64 t2 = mkc(em1.re - w.re, em1.im - w.im);
  *p2 = np_div(np_rmul(2.0, t2), w2);
  if (need3) {
    z64 w3 = np_mul_nofma(np_mul_nofma(w, w), w);
    z64 h2 = np_rdiv(w2, 2.0);
    z64 t3 = mkc(t2.re - h2.re, t2.im - h2.im);
    *p3 = np_div(np_rmul(6.0, t3), w3);
  }
}
__device__ __forceinline__ void np_psi123(z64 w, int need3, z64* p1, z64* p2, z64* p3) {
  np_psi123_e(w, np_exp(w), need3, p1, p2, p3);
}
