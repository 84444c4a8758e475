// Stepping engine for gfx950: one workgroup integrates one trajectory.
//
// A persistent kernel pulls instances from an atomic work queue.  For each
// instance the workgroup runs the reference's whole control flow on the device
// (solveras.py:336-410 step, 556-650 evolve; solvercs.py:163-279):
//   * coefficient arrays are regenerated in FP64 when h changes
//     (exact compare, etd35.py:851-852), stored in engine precision;
//   * every stage is one pass over the modes: k = E u + sum C NL is formed in
//     registers and written straight into the LDS FFT buffer, the nonlinear
//     term is evaluated in LDS (inverse FFT, pointwise product, forward FFT)
//     and the new NL array is written once;
//   * the embedded error norm (solveras.py:449-455) is a workgroup reduction
//     accumulated in FP64, and the accept/reject dt controller
//     (solveras.py:457-554) runs in FP64 on the device.
// Nonlinear terms (built in):
//   NLK_ADVECT  coef * rfft(irfft(u) * irfft(i kx u))   (KS, Burgers, KdV)
//   NLK_CUBIC   i gamma fft(|v|^2 v), v = ifft(u)        (cubic NLS)
// Workspace per instance (engine precision complex, m modes each):
//   nslot NL arrays | 2 candidate buffers | ncoef coefficient arrays.
#pragma once
#include <hip/hip_runtime.h>

#include "rks_fft.h"
#include "rks_math.h"
#include "rks_types.h"

#define NLK_ADVECT 0
#define NLK_CUBIC 1
#define NLK_NONE 2

struct rks_launch {
  const rks_tableau* tab;
  const rks_ctrl* ctl;
  rks_inst* inst;
  const rks_io* io;
  const int* order;
  int count;
  int mode;
  int* queue;
  const void* tw;  // exp(-2 pi i x / N), x < N, engine precision
  double h_mode;   // RKS_MODE_COEFFS
};

// threads per workgroup for an N-point problem
__host__ __device__ constexpr int rks_threads(int log2n, int nlk) {
  const int N = 1 << log2n;
  const int lo = (nlk == NLK_ADVECT) ? (N / 4 < 64 ? N / 4 : 64) : (N / 2 < 64 ? N / 2 : 64);
  const int t = N / 16 > lo ? N / 16 : lo;
  return t > 256 ? 256 : t;
}

__device__ __forceinline__ double nanmax(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  return a > b ? a : b;
}

template <int T>
__device__ __forceinline__ double wg_max(double v, double* red) {
  constexpr int W = T < 64 ? T : 64;
#pragma unroll
  for (int o = W / 2; o >= 1; o >>= 1) v = nanmax(v, __shfl_xor(v, o, 64));
  if (T <= 64) return v;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int i = 1; i < T / 64; ++i) r = nanmax(r, red[i]);
  __syncthreads();
  return r;
}

template <int T>
__device__ __forceinline__ void wg_sum2(double& a, double& b, double* red) {
  constexpr int W = T < 64 ? T : 64;
#pragma unroll
  for (int o = W / 2; o >= 1; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  if (T <= 64) return;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    red[2 * wid] = a;
    red[2 * wid + 1] = b;
  }
  __syncthreads();
  a = red[0];
  b = red[1];
#pragma unroll
  for (int i = 1; i < T / 64; ++i) {
    a += red[2 * i];
    b += red[2 * i + 1];
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// FP64 coefficients of one mode (etd35.py:169-288 and the other builders)
// ---------------------------------------------------------------------------
__device__ __forceinline__ z64 frac_arg(z64 z, int num, int den, int cplx) {
#pragma clang fp contract(off)
  z64 w = (num == 1) ? z : np_rmul((double)num, z);
  if (den == 1) return w;
  if (cplx) return np_rdiv(w, (double)den);
  return mkc(w.re / (double)den, w.im / (double)den);
}

template <typename R>
__device__ void coef_mode(const rks_tableau& tb, const rks_ctrl& ct, const double* __restrict__ L,
                          int L_complex, int m, int M, double h, cx<R>* __restrict__ coefs) {
#pragma clang fp contract(off)
  // ETD methods cast lin_op to complex128 (etd35.py:124); IF methods keep its dtype
  const int cplx = L_complex || tb.uses_contour;
  z64 Lm = L_complex ? mkc(L[2 * m], L[2 * m + 1]) : mkc(L[m], 0.0);
  z64 z = mkc(h * Lm.re, h * Lm.im);
  const bool small = tb.uses_contour && (hypot(z.re, z.im) < ct.modecutoff);
  z64 B[RKS_MAX_BASIS];
#pragma unroll
  for (int b = 0; b < RKS_MAX_BASIS; ++b) {
    if (b >= tb.nbasis) break;
    const rks_basis bs = tb.basis[b];
    if (bs.r == RKS_BASIS_EXP) {
      B[b] = np_exp(frac_arg(z, bs.fnum, bs.fden, cplx));
    } else if (!small) {
      z64 p1, p2, p3;
      np_psi123(frac_arg(z, bs.fnum, bs.fden, 1), bs.r == RKS_BASIS_PSI3, &p1, &p2, &p3);
      const z64 p = bs.r == RKS_BASIS_PSI1 ? p1 : (bs.r == RKS_BASIS_PSI2 ? p2 : p3);
      B[b] = np_rmul(h, p);
    } else {
      // contour mean over nodes R exp(2 pi i (j + 1/2) / N)   (etd35.py:255-264)
      const int np_ = ct.contour_points;
      const double inv_np = 1.0 / (double)np_;
      z64 acc[4] = {mkc(0.0, 0.0), mkc(0.0, 0.0), mkc(0.0, 0.0), mkc(0.0, 0.0)};
      for (int j = 0; j < np_; ++j) {
        const double th = (6.283185307179586 * (j + 0.5)) * inv_np;
        double sn, cs;
        sincos(th, &sn, &cs);
        const z64 node = mkc(ct.contour_radius * cs, ct.contour_radius * sn);
        const z64 Z = mkc(z.re + node.re, z.im + node.im);
        z64 p1, p2, p3;
        np_psi123(frac_arg(Z, bs.fnum, bs.fden, 1), bs.r == RKS_BASIS_PSI3, &p1, &p2, &p3);
        const z64 p = bs.r == RKS_BASIS_PSI1 ? p1 : (bs.r == RKS_BASIS_PSI2 ? p2 : p3);
        acc[j & 3] = mkc(acc[j & 3].re + p.re, acc[j & 3].im + p.im);
      }
      const z64 s = mkc((acc[0].re + acc[1].re) + (acc[2].re + acc[3].re),
                        (acc[0].im + acc[1].im) + (acc[2].im + acc[3].im));
      B[b] = np_rdiv(np_rmul(h, s), (double)np_);
    }
  }
  for (int c = 0; c < tb.ncoef; ++c) {
    const rks_coef cf = tb.coef[c];
    z64 acc = mkc(0.0, 0.0);
    for (int k = 0; k < cf.nb; ++k) {
      int bi = cf.basis[k];
      z64 bv = B[0];
#pragma unroll
      for (int b = 1; b < RKS_MAX_BASIS; ++b)
        if (b == bi) bv = B[b];
      const double w = cf.hmul ? cf.w[k] * h : cf.w[k];
      const z64 t = (w == 1.0) ? bv : np_rmul(w, bv);
      acc = (k == 0) ? t : mkc(acc.re + t.re, acc.im + t.im);
    }
    if (cf.outer != 1.0) acc = np_rmul(cf.outer, acc);
    if (cf.den != 1.0) {
      if (cplx) acc = np_rdiv(acc, cf.den);
      else acc = mkc(acc.re / cf.den, acc.im / cf.den);
    }
    coefs[(size_t)c * M + m] = mkc((R)acc.re, (R)acc.im);
  }
}

// ---------------------------------------------------------------------------
// one term of a stage combination / error estimate (grouping of rks_term)
// ---------------------------------------------------------------------------
template <typename R, typename SlotF>
__device__ __forceinline__ cx<R> term_value(const rks_term& tm, SlotF slotp, const cx<R>* __restrict__ coefs,
                                            size_t M, int m, double h) {
  using C = cx<R>;
  C v = slotp(tm.src)[m];
  if (tm.src2 >= 0) v = v + (R)tm.scale2 * slotp(tm.src2)[m];
  C c;
  if (tm.hmode == 0) {
    c = (tm.coef < 0) ? v : coefs[(size_t)tm.coef * M + m] * v;
    if (tm.scale != 1.0) c = (R)tm.scale * c;
  } else if (tm.hmode == 1) {
    const R hr = (R)h;
    c = (tm.coef < 0) ? hr * v : (hr * coefs[(size_t)tm.coef * M + m]) * v;
    if (tm.scale != 1.0) c = (R)tm.scale * c;
  } else {
    c = (R)((tm.scale * h) / tm.den) * v;
  }
  return c;
}

// k = E u + terms (or E u + h * sum(terms) when the stage factors h out)
template <typename R, typename SlotF>
__device__ __forceinline__ cx<R> stage_value(const rks_stage& S, const cx<R>* __restrict__ u, SlotF slotp,
                                             const cx<R>* __restrict__ coefs, size_t M, int m, double h) {
  using C = cx<R>;
  C acc = (S.ecoef < 0) ? u[m] : coefs[(size_t)S.ecoef * M + m] * u[m];
  if (S.hsum) {
    C sum = term_value<R>(S.t[0], slotp, coefs, M, m, h);
    for (int j = 1; j < S.nterm; ++j) sum = sum + term_value<R>(S.t[j], slotp, coefs, M, m, h);
    acc = acc + (R)h * sum;
  } else {
    for (int j = 0; j < S.nterm; ++j) acc = acc + term_value<R>(S.t[j], slotp, coefs, M, m, h);
  }
  return acc;
}

// embedded error estimate of the tableau (etd35.py:344, if34.py:130, if45dp.py:170-177)
template <typename R, typename SlotF>
__device__ __forceinline__ cx<R> error_value(const rks_tableau& tb, SlotF slotp, const cx<R>* __restrict__ coefs,
                                             size_t M, int m, double h) {
  using C = cx<R>;
  C ev = term_value<R>(tb.err[0], slotp, coefs, M, m, h);
  for (int j = 1; j < tb.nerr; ++j) ev = ev + term_value<R>(tb.err[j], slotp, coefs, M, m, h);
  if (tb.err_factored) {
    if (tb.err_hpow) ev = (R)h * ev;
    if (tb.err_scale != 1.0) ev = (R)tb.err_scale * ev;
    if (tb.err_coef >= 0) ev = coefs[(size_t)tb.err_coef * M + m] * ev;
  }
  return ev;
}

// ---------------------------------------------------------------------------
// nonlinear term in LDS: fill from getk(m), transform, write dst (mode-owner
// mapping m = tid + i*T is the same for fill and write-back)
// ---------------------------------------------------------------------------
template <typename R, int LOG2N, int T, int NLK, typename GetK>
__device__ __forceinline__ void eval_nl(GetK getk, cx<R>* buf, const cx<R>* __restrict__ tw,
                                        const R* __restrict__ kx, R coef, cx<R>* __restrict__ dst) {
  using C = cx<R>;
  constexpr int N = 1 << LOG2N;
  const int t = threadIdx.x;
  const R inv = (R)1 / (R)N;
  if constexpr (NLK == NLK_CUBIC) {
#pragma unroll
    for (int i = 0; i < N / T; ++i) {
      const int m = t + i * T;
      buf[m] = getk(m);
    }
    __syncthreads();
    rksfft::fft<R, N, T, true>(buf, tw, 1);
#pragma unroll
    for (int i = 0; i < N / T; ++i) {
      const int j = t + i * T;
      const C v = inv * buf[j];
      const R a = v.re * v.re + v.im * v.im;
      buf[j] = a * v;
    }
    __syncthreads();
    rksfft::fft<R, N, T, false>(buf, tw, 1);
#pragma unroll
    for (int i = 0; i < N / T; ++i) {
      const int m = t + i * T;
      const C w = buf[m];
      dst[m] = mkc(-coef * w.im, coef * w.re);
    }
    __syncthreads();
  } else {
    constexpr int M = N / 2 + 1;
    constexpr int MPT = (M + T - 1) / T;
    constexpr int H = N / 2;
#pragma unroll
    for (int i = 0; i < MPT; ++i) {
      const int m = t + i * T;
      if (m < M) {
        const C U = getk(m);
        const R k = kx[m];
        const C V = mkc(-k * U.im, k * U.re);
        if (m == 0 || m == H) {
          buf[m] = mkc(U.re, V.re);
        } else {
          buf[m] = mkc(U.re - V.im, U.im + V.re);
          buf[N - m] = mkc(U.re + V.im, V.re - U.im);
        }
      }
    }
    __syncthreads();
    rksfft::fft<R, N, T, true>(buf, tw, 1);
    C zz[H / T];
#pragma unroll
    for (int i = 0; i < H / T; ++i) {
      const int j = t + i * T;
      const C a = buf[2 * j], b = buf[2 * j + 1];
      zz[i] = mkc((a.re * inv) * (a.im * inv), (b.re * inv) * (b.im * inv));
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < H / T; ++i) buf[t + i * T] = zz[i];
    __syncthreads();
    rksfft::fft<R, H, T, false>(buf, tw, 2);
#pragma unroll
    for (int i = 0; i < MPT; ++i) {
      const int m = t + i * T;
      if (m < M) {
        const C Zk = buf[m & (H - 1)];
        const C Zc = conjc(buf[(H - m) & (H - 1)]);
        const C E = (R)0.5 * (Zk + Zc);
        const C D = Zk - Zc;
        const C O = mkc((R)0.5 * D.im, (R)-0.5 * D.re);
        const C W = E + tw[m] * O;
        dst[m] = coef * W;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// one instance: evolve / step / nl / coefficients
// ---------------------------------------------------------------------------
template <typename R, int LOG2N, int T, int NLK>
__device__ void run_instance(const rks_launch& P, int idx, cx<R>* buf, double* red) {
  using C = cx<R>;
  constexpr int N = 1 << LOG2N;
  constexpr int M = (NLK == NLK_ADVECT) ? N / 2 + 1 : N;
  constexpr int MPT = (M + T - 1) / T;
  const rks_tableau& tb = *P.tab;
  const rks_ctrl& ct = *P.ctl;
  rks_inst* in = P.inst + idx;
  const rks_io io = P.io[idx];
  const int t = threadIdx.x;
  const C* tw = (const C*)P.tw;
  const R* kx = (const R*)in->kx;
  const R nlc = (R)in->nl_coef;
  const double* Ld = in->L;
  const int Lc = in->L_complex;
  C* ws = (C*)in->ws;
  const int nslot = tb.nslot;
  C* slot_base = ws;
  C* candb[2] = {ws + (size_t)nslot * M, ws + (size_t)(nslot + 1) * M};
  C* coefs = ws + (size_t)(nslot + 2) * M;

  if (P.mode == RKS_MODE_NL) {
    const C* u = (const C*)io.u_in;
    eval_nl<R, LOG2N, T, NLK>([&](int m) { return u[m]; }, buf, tw, kx, nlc, (C*)io.u_out);
    return;
  }
  if (P.mode == RKS_MODE_COEFFS) {
    for (int i = 0; i < MPT; ++i) {
      const int m = t + i * T;
      if (m < M) coef_mode<R>(tb, ct, Ld, Lc, m, M, P.h_mode, coefs);
    }
    return;
  }

  // ---- load persistent state (uniform across the workgroup) ----
  double tc = in->tc, tf = in->tf, h = in->h, h_coeff = in->h_coeff, h_next = in->h_next;
  double h_used = in->h_used;
  int started = in->started, accept = in->accept, steps = in->steps, attempts = in->attempts;
  int rejects = in->rejects, status = in->status, ubuf = in->ubuf;
  int log_count = in->log_count, snap_count = in->snap_count;
  int nl_calls = in->nl_calls, coef_updates = in->coef_updates;
  int smap[RKS_MAX_SLOTS];
#pragma unroll
  for (int s = 0; s < RKS_MAX_SLOTS; ++s) smap[s] = in->slot_map[s];
  const int store_freq = in->store_freq, snap_cap = in->snap_cap, log_cap = in->log_cap;

  if (io.reset) {
    started = 0;
    accept = 0;
    h_coeff = __builtin_nan("");
#pragma unroll
    for (int s = 0; s < RKS_MAX_SLOTS; ++s) smap[s] = s;
  }
  const bool evolve = (P.mode == RKS_MODE_EVOLVE);
  const C* uin = (const C*)io.u_in;
  if (evolve && !io.resume) {
    tc = io.t0;
    tf = io.tf;
    h = io.h_in;
    steps = attempts = rejects = 0;
    log_count = snap_count = 0;
    nl_calls = coef_updates = 0;
    ubuf = -1;
    status = RKS_ST_RUNNING;
    if (tb.adaptive && tc + h > tf) h = tf - tc;  // solveras.py:620-621
    if (store_freq > 0) {
      for (int i = 0; i < MPT; ++i) {
        const int m = t + i * T;
        if (m < M) ((C*)in->snap)[m] = uin[m];
      }
      if (t == 0) in->snap_t[0] = tc;
      snap_count = 1;
    }
  } else if (evolve) {
    status = RKS_ST_RUNNING;
  } else {
    h = io.h_in;
    status = RKS_ST_RUNNING;
    attempts = rejects = 0;
  }

  auto slot = [&](int s) -> C* { return slot_base + (size_t)smap[s] * M; };
  auto cur_u = [&]() -> const C* { return ubuf < 0 ? uin : (const C*)candb[ubuf]; };

  bool done = false;
  while (!done) {
    if (evolve) {
      if (!(tc < tf)) {
        status = RKS_ST_OK;
        break;
      }
      if (store_freq > 0 && snap_count >= snap_cap) {
        status = RKS_ST_PAUSED;
        break;
      }
    }
    const C* u = evolve ? cur_u() : uin;
    C* cand = evolve ? candb[ubuf == 0 ? 1 : 0] : (C*)io.u_out;
    int numloops = 0;
    double h_suggest = h;
    // ---------------- attempt loop (solveras.py:384-410) ----------------
    for (;;) {
      if (h != h_coeff) {
        for (int i = 0; i < MPT; ++i) {
          const int m = t + i * T;
          if (m < M) coef_mode<R>(tb, ct, Ld, Lc, m, M, h, coefs);
        }
        h_coeff = h;
        ++coef_updates;
      }
      if (!started) {
        eval_nl<R, LOG2N, T, NLK>([&](int m) { return u[m]; }, buf, tw, kx, nlc, slot(0));
        ++nl_calls;
        started = 1;
      } else if (accept || !tb.adaptive) {
        if (tb.nl1_policy == RKS_NL1_RECOMPUTE) {
          eval_nl<R, LOG2N, T, NLK>([&](int m) { return u[m]; }, buf, tw, kx, nlc, slot(0));
          ++nl_calls;
        } else {
          const int a = smap[0];
          smap[0] = smap[tb.fsal_slot];
          smap[tb.fsal_slot] = a;
        }
      }
      for (int s = 0; s < tb.nstage; ++s) {
        const rks_stage& S = tb.st[s];
        const bool last = (s == tb.nstage - 1);
        auto getk = [&](int m) -> C {
          const C acc = stage_value<R>(S, u, slot, coefs, M, m, h);
          if (last) cand[m] = acc;
          return acc;
        };
        if (S.has_nl) {
          eval_nl<R, LOG2N, T, NLK>(getk, buf, tw, kx, nlc, slot(S.nl_slot));
          ++nl_calls;
        } else {
          for (int i = 0; i < MPT; ++i) {
            const int m = t + i * T;
            if (m < M) (void)getk(m);
          }
        }
      }
      if (!tb.adaptive) {
        h_suggest = h;
        accept = 1;
        break;
      }
      // ---------------- embedded error norm (solveras.py:449-455) ----------------
      double kmag[MPT], k2[MPT], e2[MPT];
      double mx = 0.0;
#pragma unroll
      for (int i = 0; i < MPT; ++i) {
        const int m = t + i * T;
        kmag[i] = 0.0;
        k2[i] = 0.0;
        e2[i] = 0.0;
        if (m < M) {
          const C kv = cand[m];
          const C ev = error_value<R>(tb, slot, coefs, M, m, h);
          const double kr = (double)kv.re, ki = (double)kv.im;
          const double er = (double)ev.re, ei = (double)ev.im;
          kmag[i] = hypot(kr, ki);
          k2[i] = kr * kr + ki * ki;
          e2[i] = er * er + ei * ei;
          mx = (i == 0) ? kmag[i] : nanmax(mx, kmag[i]);
        }
      }
      if (t >= M) mx = 0.0;
      mx = wg_max<T>(mx, red);
      double su = 0.0, se = 0.0;
#pragma unroll
      for (int i = 0; i < MPT; ++i) {
        const int m = t + i * T;
        if (m < M && (kmag[i] / mx > ct.adapt_cutoff)) {
          su += k2[i];
          se += e2[i];
        }
      }
      wg_sum2<T>(su, se, red);
      const double tol = ct.epsilon * sqrt(su);
      const double sv = ct.safety_f * pow(tol / sqrt(se), 1.0 / (double)tb.q);
      ++attempts;
      if (t == 0 && log_count < log_cap) {
        in->log[2 * log_count] = h;
        in->log[2 * log_count + 1] = sv;
      }
      ++log_count;
      if (isinf(sv) || isnan(sv) || sv < 1.0) {
        // reject (solveras.py:490-506)
        accept = 0;
        ++rejects;
        if (isinf(sv) || isnan(sv)) {
          h = ct.min_s * h;
        } else {
          double sc = sv > ct.min_s ? sv : ct.min_s;
          sc = sc < ct.decr_f ? sc : ct.decr_f;
          h = sc * h;
        }
        ++numloops;
        if (numloops > ct.max_loops) {
          status = RKS_ST_MAXLOOPS;
          done = true;
          break;
        }
        if (h < ct.minh) {
          status = RKS_ST_MINSTEP;
          done = true;
          break;
        }
        continue;
      }
      // accept (solveras.py:543-554)
      accept = 1;
      const double sc = sv < ct.max_s ? sv : ct.max_s;
      h_suggest = (sc > ct.incr_f) ? sc * h : h;
      break;
    }
    if (done) break;
    // ---------------- accepted step ----------------
    h_used = h;
    h_next = h_suggest;
    if (!evolve) {
      status = RKS_ST_OK;
      ++steps;
      break;
    }
    ubuf = (ubuf == 0) ? 1 : 0;
    const double tprev = tc;
    tc += h;
    ++steps;
    if (!(tc > tprev)) {
      // h below the resolution of tc: the reference would loop forever here;
      // stop with the minimum-step status instead of spinning on the device
      status = RKS_ST_MINSTEP;
      break;
    }
    if (tb.adaptive) h = (tc + h_suggest > tf) ? tf - tc : h_suggest;  // solveras.py:638-641
    if (store_freq > 0 && (steps % store_freq) == 0) {
      C* dst = (C*)in->snap + (size_t)snap_count * M;
      const C* src = cur_u();
      for (int i = 0; i < MPT; ++i) {
        const int m = t + i * T;
        if (m < M) dst[m] = src[m];
      }
      if (t == 0) in->snap_t[snap_count] = tc;
      ++snap_count;
    }
  }
  if (evolve && status != RKS_ST_PAUSED) {
    const C* src = cur_u();
    C* dst = (C*)io.u_out;
    for (int i = 0; i < MPT; ++i) {
      const int m = t + i * T;
      if (m < M) dst[m] = src[m];
    }
  }
  __syncthreads();
  if (t == 0) {
    in->tc = tc;
    in->tf = tf;
    in->h = h;
    in->h_coeff = h_coeff;
    in->h_next = h_next;
    in->h_used = h_used;
    in->started = started;
    in->accept = accept;
    in->steps = steps;
    in->attempts = attempts;
    in->rejects = rejects;
    in->status = status;
    in->ubuf = ubuf;
    in->log_count = log_count;
    in->snap_count = snap_count;
    in->nl_calls = nl_calls;
    in->coef_updates = coef_updates;
#pragma unroll
    for (int s = 0; s < RKS_MAX_SLOTS; ++s) in->slot_map[s] = smap[s];
  }
}

template <typename R, int LOG2N, int NLK>
__global__ __launch_bounds__(rks_threads(LOG2N, NLK)) void rks_kernel(rks_launch P) {
  constexpr int T = rks_threads(LOG2N, NLK);
  extern __shared__ __align__(16) unsigned char rks_smem[];
  __shared__ double red[2 * (T / 64 + 1)];
  __shared__ int s_item;
  cx<R>* buf = (cx<R>*)rks_smem;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(P.queue, 1);
    __syncthreads();
    const int item = s_item;
    __syncthreads();
    if (item >= P.count) break;
    run_instance<R, LOG2N, T, NLK>(P, P.order[item], buf, red);
    __syncthreads();
  }
}

