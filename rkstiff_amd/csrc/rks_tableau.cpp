// Method tableaux: the seven diagonal schemes of the reference as data.
//
// Each entry restates one strategy class of the reference:
//   ETD35  etd35.py:169-288 (coefficients), 301-345 (stages), 880-889 (q=4)
//   ETD34  etd34.py:96-149, 162-191, 611-621
//   ETD4   etd4.py:87-139, 152-175 (constant step)
//   ETD5   etd5.py:115-203, 219-261 (constant step)
//   IF34   if34.py:81-93, 99-131, 336-338
//   IF45DP if45dp.py:112-181, 183-238, 240-249 (q=5)
//   IF4    if4.py:73-83, 96-122 (constant step)
// Coefficients are rational combinations of basis values (h psi_r(f z) or
// exp(f z)); stage k = E u + sum scale h^p C NL.  Grouped reference terms such
// as a52 (NL2 - NL3) become two terms sharing one coefficient array.
#include <cstring>

#include "rks_types.h"

namespace {

struct Builder {
  rks_tableau t;
  Builder(int method, int adaptive, int q, int policy, int nslot, int fsal_slot, int contour) {
    std::memset(&t, 0, sizeof(t));
    t.method = method;
    t.adaptive = adaptive;
    t.q = q;
    t.nl1_policy = policy;
    t.nslot = nslot;
    t.fsal_slot = fsal_slot;
    t.uses_contour = contour;
  }
  // basis handle = fraction index * 4 + kind
  int basis(int r, int fnum, int fden) {
    int f = 0;
    for (; f < t.nfrac; ++f)
      if (t.frac[f].fnum == fnum && t.frac[f].fden == fden) break;
    if (f == t.nfrac) {
      t.frac[f].fnum = fnum;
      t.frac[f].fden = fden;
      t.frac[f].mask = 0;
      ++t.nfrac;
    }
    t.frac[f].mask |= 1 << r;
    return 4 * f + r;
  }
  // c = (outer * sum w_k B_k) / den  (psi bases already carry h); one fraction per recipe
  int coef(int nb, const int* bs, const double* ws, double den, int hmul = 0, double outer = 1.0) {
    rks_coef& c = t.coef[t.ncoef];
    c.nb = nb;
    c.hmul = hmul;
    c.outer = outer;
    c.frac = bs[0] / 4;
    for (int i = 0; i < nb; ++i) {
      c.kind[i] = bs[i] % 4;
      c.w[i] = ws[i];
    }
    c.den = den;
    return t.ncoef++;
  }
  int coef1(int b, double w = 1.0, double den = 1.0, int hmul = 0) {
    int bs[1] = {b};
    double ws[1] = {w};
    return coef(1, bs, ws, den, hmul);
  }
  rks_stage& stage(int ecoef, int has_nl, int nl_slot) {
    rks_stage& s = t.st[t.nstage++];
    s.ecoef = ecoef;
    s.has_nl = has_nl;
    s.nl_slot = nl_slot;
    return s;
  }
  static rks_term mk(int src, int coef, double scale, int hmode, int src2, double scale2, double den) {
    rks_term x;
    x.src = src;
    x.src2 = src2;
    x.coef = coef;
    x.hmode = hmode;
    x.scale = scale;
    x.scale2 = scale2;
    x.den = den;
    return x;
  }
  // scale * (C * NL[src])
  static void term(rks_stage& s, int src, int coef, double scale = 1.0) {
    s.t[s.nterm++] = mk(src, coef, scale, 0, -1, 0.0, 1.0);
  }
  // C * (NL[src] + scale2 * NL[src2])
  static void pair(rks_stage& s, int src, int src2, double scale2, int coef) {
    s.t[s.nterm++] = mk(src, coef, 1.0, 0, src2, scale2, 1.0);
  }
  // ((h * C) * NL[src]) * scale
  static void hterm(rks_stage& s, int src, int coef, double scale) {
    s.t[s.nterm++] = mk(src, coef, scale, 1, -1, 0.0, 1.0);
  }
  // ((num * h) / den) * NL[src]
  static void hscalar(rks_stage& s, int src, double num, double den) {
    s.t[s.nterm++] = mk(src, RKS_COEF_ONE, num, 2, -1, 0.0, den);
  }
  void factored_err(int coef, double scale, int hpow) {
    t.err_factored = 1;
    t.err_coef = coef;
    t.err_scale = scale;
    t.err_hpow = hpow;
  }
  void err(const rks_term& x) { t.err[t.nerr++] = x; }
  // propagators (coefficients multiplying u) keep the state precision
  rks_tableau finalize() {
    for (int c = 0; c < t.ncoef; ++c) t.coef_rs[c] = 0;
    for (int s = 0; s < t.nstage; ++s)
      if (t.st[s].ecoef >= 0) t.coef_rs[t.st[s].ecoef] = 1;
    int off = 0;
    for (int c = 0; c < t.ncoef; ++c) {
      t.coef_off[c] = off;
      off += 16;
    }
    return t;
  }
};

// ETD35 / ETD5 share the fifth-order coefficient rows (etd35.py:213-237).
struct Fifth {
  int EL14, EL12, EL34, EL, a21, a31, a32, a41, a43, a51, a52, a54, a61, a62, a63, a65, a71, a73, a74,
      a75, a76;
};

Fifth fifth_coeffs(Builder& B) {
  Fifth F;
  const int p1_14 = B.basis(RKS_BASIS_PSI1, 1, 4), p2_14 = B.basis(RKS_BASIS_PSI2, 1, 4);
  const int p1_12 = B.basis(RKS_BASIS_PSI1, 1, 2), p2_12 = B.basis(RKS_BASIS_PSI2, 1, 2);
  const int p1_34 = B.basis(RKS_BASIS_PSI1, 3, 4), p2_34 = B.basis(RKS_BASIS_PSI2, 3, 4);
  const int p1 = B.basis(RKS_BASIS_PSI1, 1, 1), p2 = B.basis(RKS_BASIS_PSI2, 1, 1);
  const int p3 = B.basis(RKS_BASIS_PSI3, 1, 1);
  F.EL14 = B.coef1(B.basis(RKS_BASIS_EXP, 1, 4));
  F.EL12 = B.coef1(B.basis(RKS_BASIS_EXP, 1, 2));
  F.EL34 = B.coef1(B.basis(RKS_BASIS_EXP, 3, 4));
  F.EL = B.coef1(B.basis(RKS_BASIS_EXP, 1, 1));
  F.a21 = B.coef1(p1_14, 1.0, 4.0);
  {
    int b[2] = {p1_14, p2_14};
    double w[2] = {1.0, -0.5};
    F.a31 = B.coef(2, b, w, 4.0);
  }
  F.a32 = B.coef1(p2_14, 1.0, 8.0);
  {
    int b[2] = {p1_12, p2_12};
    double w[2] = {1.0, -1.0};
    F.a41 = B.coef(2, b, w, 2.0);
  }
  F.a43 = B.coef1(p2_12, 1.0, 2.0);
  {
    // 3 (p1_34 - 3 p2_34 / 4) / 4
    int b[2] = {p1_34, p2_34};
    double w[2] = {1.0, -0.75};
    F.a51 = B.coef(2, b, w, 4.0, 0, 3.0);
  }
  F.a52 = B.coef1(p1_34, -3.0, 8.0);
  F.a54 = B.coef1(p2_34, 9.0, 16.0);
  {
    int b[2] = {p1, p2};
    double w[2] = {-77.0, 59.0};
    F.a61 = B.coef(2, b, w, 42.0);
  }
  F.a62 = B.coef1(p1, 8.0, 7.0);
  {
    int b[2] = {p1, p2};
    double w[2] = {111.0, -87.0};
    F.a63 = B.coef(2, b, w, 28.0);
  }
  {
    int b[2] = {p1, p2};
    double w[2] = {-47.0, 143.0};
    F.a65 = B.coef(2, b, w, 84.0);
  }
  int b3[3] = {p1, p2, p3};
  {
    double w[3] = {257.0, -497.0, 270.0};
    F.a71 = B.coef(3, b3, w, 2700.0, 0, 7.0);
  }
  {
    // the reference notes the paper's 497 is 467 here (etd35.py:233-234)
    double w[3] = {1097.0, -467.0, -150.0};
    F.a73 = B.coef(3, b3, w, 1350.0);
  }
  {
    double w[3] = {-49.0, 199.0, -135.0};
    F.a74 = B.coef(3, b3, w, 225.0, 0, 2.0);
  }
  {
    double w[3] = {-313.0, 883.0, -90.0};
    F.a75 = B.coef(3, b3, w, 1350.0);
  }
  {
    double w[3] = {509.0, -2129.0, 1830.0};
    F.a76 = B.coef(3, b3, w, 2700.0);
  }
  return F;
}

// stages 2..6 shared by ETD35 and ETD5 (etd35.py:320-334, etd5.py:236-250)
void fifth_stages(Builder& B, const Fifth& F) {
  rks_stage& s2 = B.stage(F.EL14, 1, 1);
  Builder::term(s2, 0, F.a21);
  rks_stage& s3 = B.stage(F.EL14, 1, 2);
  Builder::term(s3, 0, F.a31);
  Builder::term(s3, 1, F.a32);
  rks_stage& s4 = B.stage(F.EL12, 1, 3);
  Builder::term(s4, 0, F.a41);
  Builder::term(s4, 2, F.a43);
  rks_stage& s5 = B.stage(F.EL34, 1, 4);
  Builder::term(s5, 0, F.a51);
  Builder::pair(s5, 1, 2, -1.0, F.a52);  // a52 (NL2 - NL3)
  Builder::term(s5, 3, F.a54);
  rks_stage& s6 = B.stage(F.EL, 1, 5);
  Builder::term(s6, 0, F.a61);
  Builder::pair(s6, 1, 3, -1.5, F.a62);  // a62 (NL2 - 3 NL4 / 2)
  Builder::term(s6, 2, F.a63);
  Builder::term(s6, 4, F.a65);
}

rks_tableau make_etd35() {
  Builder B(0, 1, 4, RKS_NL1_RECOMPUTE, 6, 0, 1);
  Fifth F = fifth_coeffs(B);
  fifth_stages(B, F);
  rks_stage& s7 = B.stage(F.EL, 0, 0);
  Builder::term(s7, 0, F.a71);
  Builder::term(s7, 2, F.a73);
  Builder::term(s7, 3, F.a74);
  Builder::term(s7, 4, F.a75);
  Builder::term(s7, 5, F.a76);
  // err = a75 (-NL1 + 4 NL3 - 6 NL4 + 4 NL5 - NL6)   (etd35.py:344), summed first
  B.factored_err(F.a75, 1.0, 0);
  B.err(Builder::mk(0, RKS_COEF_ONE, -1.0, 0, -1, 0.0, 1.0));
  B.err(Builder::mk(2, RKS_COEF_ONE, 4.0, 0, -1, 0.0, 1.0));
  B.err(Builder::mk(3, RKS_COEF_ONE, -6.0, 0, -1, 0.0, 1.0));
  B.err(Builder::mk(4, RKS_COEF_ONE, 4.0, 0, -1, 0.0, 1.0));
  B.err(Builder::mk(5, RKS_COEF_ONE, -1.0, 0, -1, 0.0, 1.0));
  return B.finalize();
}

rks_tableau make_etd5() {
  Builder B(3, 0, 5, RKS_NL1_FSAL, 7, 6, 1);
  Fifth F = fifth_coeffs(B);
  fifth_stages(B, F);
  rks_stage& s7 = B.stage(F.EL, 1, 6);  // FSAL: NL1 = nl(k)  (etd5.py:260)
  Builder::term(s7, 0, F.a71);            // b1
  Builder::term(s7, 2, F.a73);            // b3
  Builder::term(s7, 3, F.a74);            // b4
  Builder::term(s7, 4, F.a75);            // b5
  Builder::term(s7, 5, F.a76);            // b6
  return B.finalize();
}

struct Fourth {
  int EL, EL2, a21, a31, a32, a41, a43, a51, a52, a54;
};

Fourth fourth_coeffs(Builder& B) {
  Fourth F;
  const int p1_12 = B.basis(RKS_BASIS_PSI1, 1, 2), p2_12 = B.basis(RKS_BASIS_PSI2, 1, 2);
  const int p1 = B.basis(RKS_BASIS_PSI1, 1, 1), p2 = B.basis(RKS_BASIS_PSI2, 1, 1);
  const int p3 = B.basis(RKS_BASIS_PSI3, 1, 1);
  F.EL = B.coef1(B.basis(RKS_BASIS_EXP, 1, 1));
  F.EL2 = B.coef1(B.basis(RKS_BASIS_EXP, 1, 2));
  F.a21 = B.coef1(p1_12, 0.5);
  {
    int b[2] = {p1_12, p2_12};
    double w[2] = {1.0, -1.0};
    F.a31 = B.coef(2, b, w, 1.0, 0, 0.5);
  }
  F.a32 = B.coef1(p2_12, 0.5);
  {
    int b[2] = {p1, p2};
    double w[2] = {1.0, -1.0};
    F.a41 = B.coef(2, b, w, 1.0);
  }
  F.a43 = B.coef1(p2);
  int b3[3] = {p1, p2, p3};
  {
    double w[3] = {1.0, -(3.0 / 2), (2.0 / 3)};
    F.a51 = B.coef(3, b3, w, 1.0);
  }
  {
    int b[2] = {p2, p3};
    double w[2] = {1.0, -(2.0 / 3)};
    F.a52 = B.coef(2, b, w, 1.0);
  }
  {
    int b[2] = {p2, p3};
    double w[2] = {-(1.0 / 2), (2.0 / 3)};
    F.a54 = B.coef(2, b, w, 1.0);
  }
  return F;
}

// etd34.py:182-189 / etd4.py:167-174
void fourth_stages(Builder& B, const Fourth& F, int final_nl) {
  rks_stage& s2 = B.stage(F.EL2, 1, 1);
  Builder::term(s2, 0, F.a21);
  rks_stage& s3 = B.stage(F.EL2, 1, 2);
  Builder::term(s3, 0, F.a31);
  Builder::term(s3, 1, F.a32);
  rks_stage& s4 = B.stage(F.EL, 1, 3);
  Builder::term(s4, 0, F.a41);
  Builder::term(s4, 2, F.a43);
  rks_stage& s5 = B.stage(F.EL, final_nl, 4);
  Builder::term(s5, 0, F.a51);
  Builder::pair(s5, 1, 2, 1.0, F.a52);  // a52 (NL2 + NL3)
  Builder::term(s5, 3, F.a54);
}

rks_tableau make_etd34() {
  Builder B(1, 1, 4, RKS_NL1_FSAL, 5, 4, 1);
  Fourth F = fourth_coeffs(B);
  fourth_stages(B, F, 1);
  B.factored_err(F.a54, 1.0, 0);  // err = a54 (NL4 - NL5)  (etd34.py:190)
  B.err(Builder::mk(3, RKS_COEF_ONE, 1.0, 0, 4, -1.0, 1.0));
  return B.finalize();
}

rks_tableau make_etd4() {
  Builder B(2, 0, 4, RKS_NL1_FSAL, 5, 4, 1);
  Fourth F = fourth_coeffs(B);
  fourth_stages(B, F, 1);
  return B.finalize();
}

// IF34 / IF4: integrating-factor RK4 (if34.py:120-130, if4.py:112-121)
rks_tableau make_if4family(int method, int adaptive) {
  Builder B(method, adaptive, 4, RKS_NL1_FSAL, 5, 4, 0);
  const int EL = B.coef1(B.basis(RKS_BASIS_EXP, 1, 1));
  const int EL2 = B.coef1(B.basis(RKS_BASIS_EXP, 1, 2));
  rks_stage& s2 = B.stage(EL2, 1, 1);
  Builder::hterm(s2, 0, EL2, 0.5);            // h * EL2 * NL1 / 2.0
  rks_stage& s3 = B.stage(EL2, 1, 2);
  Builder::hterm(s3, 1, RKS_COEF_ONE, 0.5);   // h * NL2 / 2.0
  rks_stage& s4 = B.stage(EL, 1, 3);
  Builder::hterm(s4, 2, EL2, 1.0);            // h * EL2 * NL3
  rks_stage& s5 = B.stage(EL, 1, 4);
  s5.hsum = 1;                                // h * (EL NL1/6 + EL2 NL2/3 + EL2 NL3/3 + NL4/6)
  Builder::term(s5, 0, EL, 1.0 / 6.0);
  Builder::term(s5, 1, EL2, 1.0 / 3.0);
  Builder::term(s5, 2, EL2, 1.0 / 3.0);
  Builder::term(s5, 3, RKS_COEF_ONE, 1.0 / 6.0);
  if (adaptive) {
    B.factored_err(RKS_COEF_ONE, 1.0 / 6.0, 1);  // err = h (NL4 - NL5) / 6.0
    B.err(Builder::mk(3, RKS_COEF_ONE, 1.0, 0, 4, -1.0, 1.0));
  }
  return B.finalize();
}

// IF45DP: Dormand-Prince in the integrating-factor frame (if45dp.py:140-238)
rks_tableau make_if45dp() {
  Builder B(5, 1, 5, RKS_NL1_FSAL, 7, 6, 0);
  auto ex = [&](int a, int b) { return B.basis(RKS_BASIS_EXP, a, b); };
  const int EL15 = B.coef1(ex(1, 5)), EL310 = B.coef1(ex(3, 10)), EL45 = B.coef1(ex(4, 5));
  const int EL89 = B.coef1(ex(8, 9)), EL = B.coef1(ex(1, 1));
  auto hc = [&](int b, double num, double den) { return B.coef1(b, num, den, 1); };
  const int a21 = hc(ex(1, 5), 1.0, 5.0);
  const int a31 = hc(ex(3, 10), 3.0, 40.0);
  const int a32 = hc(ex(1, 10), 9.0, 40.0);
  const int a41 = hc(ex(4, 5), 44.0, 45.0);
  const int a42 = hc(ex(3, 5), -56.0, 15.0);
  const int a43 = hc(ex(1, 2), 32.0, 9.0);
  const int a51 = hc(ex(8, 9), 19372.0, 6561.0);
  const int a52 = hc(ex(31, 45), -25360.0, 2187.0);
  const int a53 = hc(ex(53, 90), 64448.0, 6561.0);
  const int a54 = hc(ex(4, 45), -212.0, 729.0);
  const int a61 = hc(ex(1, 1), 9017.0, 3168.0);
  const int a62 = hc(ex(4, 5), -355.0, 33.0);
  const int a63 = hc(ex(7, 10), 46732.0, 5247.0);
  const int a64 = hc(ex(1, 5), 49.0, 176.0);
  const int a65 = hc(ex(1, 9), -5103.0, 18656.0);
  const int a71 = hc(ex(1, 1), 35.0, 384.0);
  const int a73 = hc(ex(7, 10), 500.0, 1113.0);
  const int a74 = hc(ex(1, 5), 125.0, 192.0);
  const int a75 = hc(ex(1, 9), -2187.0, 6784.0);
  const int r1 = hc(ex(1, 1), 71.0, 57600.0);
  const int r3 = hc(ex(7, 10), -71.0, 16695.0);
  const int r4 = hc(ex(1, 5), 17.0, 1920.0);
  const int r5 = hc(ex(1, 9), -17253.0, 339200.0);
  rks_stage& s2 = B.stage(EL15, 1, 1);
  Builder::term(s2, 0, a21);
  rks_stage& s3 = B.stage(EL310, 1, 2);
  Builder::term(s3, 0, a31);
  Builder::term(s3, 1, a32);
  rks_stage& s4 = B.stage(EL45, 1, 3);
  Builder::term(s4, 0, a41);
  Builder::term(s4, 1, a42);
  Builder::term(s4, 2, a43);
  rks_stage& s5 = B.stage(EL89, 1, 4);
  Builder::term(s5, 0, a51);
  Builder::term(s5, 1, a52);
  Builder::term(s5, 2, a53);
  Builder::term(s5, 3, a54);
  rks_stage& s6 = B.stage(EL, 1, 5);
  Builder::term(s6, 0, a61);
  Builder::term(s6, 1, a62);
  Builder::term(s6, 2, a63);
  Builder::term(s6, 3, a64);
  Builder::term(s6, 4, a65);
  rks_stage& s7 = B.stage(EL, 1, 6);
  Builder::term(s7, 0, a71);
  Builder::term(s7, 2, a73);
  Builder::term(s7, 3, a74);
  Builder::term(s7, 4, a75);
  Builder::hscalar(s7, 5, 11.0, 84.0);  // a76 = 11 * h / 84.0
  B.err(Builder::mk(0, r1, 1.0, 0, -1, 0.0, 1.0));
  B.err(Builder::mk(2, r3, 1.0, 0, -1, 0.0, 1.0));
  B.err(Builder::mk(3, r4, 1.0, 0, -1, 0.0, 1.0));
  B.err(Builder::mk(4, r5, 1.0, 0, -1, 0.0, 1.0));
  B.err(Builder::mk(5, RKS_COEF_ONE, 22.0, 2, -1, 0.0, 525.0));  // r6 = 22 * h / 525.0
  B.err(Builder::mk(6, RKS_COEF_ONE, -1.0, 2, -1, 0.0, 40.0));   // r7 = -h / 40.0
  return B.finalize();
}

}  // namespace

// method ids: 0 ETD35, 1 ETD34, 2 ETD4, 3 ETD5, 4 IF34, 5 IF45DP, 6 IF4
int rks_build_tableau(int method, rks_tableau* out) {
  switch (method) {
    case 0: *out = make_etd35(); return 0;
    case 1: *out = make_etd34(); return 0;
    case 2: *out = make_etd4(); return 0;
    case 3: *out = make_etd5(); return 0;
    case 4: *out = make_if4family(4, 1); return 0;
    case 5: *out = make_if45dp(); return 0;
    case 6: *out = make_if4family(6, 0); return 0;
    default: return 1;
  }
}
