// Three-dimensional non-symmetric cones: the exponential cone and the power cone
// (src/IPM/src/cone/tdc/{tdc,exp,pow,utils}.jl), one thread per cone.
//
// Shared by the device cone kernels and the host starting point, so every function is
// __host__ __device__. Scaling is Tunçel's: M = d dᵀ/⟨p,d⟩ + δd δdᵀ/⟨δp,δd⟩ + t z zᵀ
// (δp = p − μ p*, δd = d − μ d*, z = p × p* normalised), falling back to μ f''(p) near the
// central path. The barrier's Hessian is held as a pivoted analytic factor P f''(p) Pᵀ = L Lᵀ
// with pivot order (3, 1, 2). Determinants are evaluated in double-double so that they keep
// full relative accuracy near the cone boundary.
//
// Per-cone cache (19 doubles, tdc.jl:142-163): L (9, column-major), d* = −f'(p) (3),
// seed of the shadow solve (1), determinant-chain scalars d1..d4 (4), line-search values
// h_prim(0), h_dual(0) (2).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "kinds.h"

namespace mb {
namespace tdc {

#define TDC_HD __host__ __device__ __forceinline__

constexpr int CACHE = 19;
constexpr int C_L = 0, C_SD = 9, C_SEED = 12, C_D1 = 13, C_D2 = 14, C_D3 = 15, C_D4 = 16, C_H0P = 17, C_H0D = 18;
constexpr double EPS = 2.220446049250313e-16;
constexpr double TINY = 2.2250738585072014e-308;  // smallest normal
constexpr double NB_BAND = 65536.0 * EPS;          // noise band of the log-form membership functions
constexpr double NFK = 8.0;                        // line-search stop: NFK band widths

// ------------------------------------------------------------------------------------------
// double-double arithmetic (error-free transformations: src/IPM/src/utils.jl:27-38, 125-129)
// ------------------------------------------------------------------------------------------

struct dd {
  double h, l;
};

TDC_HD void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
TDC_HD void two_prod(double a, double b, double& p, double& e) {
  p = a * b;
  e = fma(a, b, -p);
}
TDC_HD dd quick(double s, double e) {  // |s| ≥ |e|
  double h = s + e;
  return dd{h, e - (h - s)};
}
TDC_HD dd dd_add(dd a, dd b) {
  double s, e, t, f;
  two_sum(a.h, b.h, s, e);
  two_sum(a.l, b.l, t, f);
  e += t;
  dd r = quick(s, e);
  return quick(r.h, r.l + f);
}
TDC_HD dd dd_neg(dd a) { return dd{-a.h, -a.l}; }
TDC_HD dd dd_mul(dd a, dd b) {
  double p, e;
  two_prod(a.h, b.h, p, e);
  e += a.h * b.l + a.l * b.h;
  return quick(p, e);
}
TDC_HD dd dd_scale(dd a, double k) {  // exact for k a power of two
  return dd{a.h * k, a.l * k};
}
// (ah + al) / (bh + bl) (twodiv, utils.jl:125-129)
TDC_HD dd twodiv(double ah, double al, double bh, double bl) {
  double q1 = ah / bh;
  double r = fma(-q1, bh, ah) + al - q1 * bl;
  dd q;
  two_sum(q1, r / bh, q.h, q.l);
  return q;
}
TDC_HD dd dd_div(dd a, dd b) { return twodiv(a.h, a.l, b.h, b.l); }

constexpr double LN2_H = 0.6931471805599453, LN2_L = 2.3190468138462996e-17;
constexpr double LOG2E_H = 1.4426950408889634, LOG2E_L = 2.0355273740931033e-17;

// atanh(s)/s = Σ_j s^{2j}/(2j+1) in double-double for |s| ≤ 0.1716 (22 terms: the tail is
// below 2⁻¹¹⁰ relative)
TDC_HD dd atanh_over_s(dd s) {
  const dd t = dd_mul(s, s);
  dd acc = twodiv(1.0, 0.0, 43.0, 0.0);
  for (int j = 20; j >= 0; --j) acc = dd_add(dd_mul(acc, t), twodiv(1.0, 0.0, 2.0 * j + 1.0, 0.0));
  return acc;
}

// log(x1/x2) for normal x1, x2 > 0 as a double-double with full relative accuracy at every
// scale (the reference's twolograt, utils.jl:84-121, computes the same two-word value from
// DoubleFloats' log₂ tables; here: the atanh series log m = 2 atanh((m−1)/(m+1)) after
// reducing the ratio to m ∈ [1/√2, √2), and for ratios near 1 the exact difference x1 − x2
// (Sterbenz) so that tiny logs stay relative-accurate)
TDC_HD dd log_ratio(double x1, double x2) {
  const double rho = x1 / x2;
  int e2;
  if (rho > 0.9375 && rho < 1.0625) {
    (void)frexp(x2, &e2);
    const double a = ldexp(x1, -e2), b = ldexp(x2, -e2);
    const dd u = twodiv(a - b, 0.0, b, 0.0);  // a − b exact
    const dd s = dd_div(u, dd_add(u, dd{2.0, 0.0}));
    return dd_scale(dd_mul(s, atanh_over_s(s)), 2.0);
  }
  int e1;
  const double m1 = frexp(x1, &e1), m2 = frexp(x2, &e2);
  int k = e1 - e2;
  dd q = twodiv(m1, 0.0, m2, 0.0);  // ∈ (1/2, 2)
  if (q.h > 1.4142135623730951) {
    q = dd_scale(q, 0.5);
    k += 1;
  } else if (q.h < 0.7071067811865476) {
    q = dd_scale(q, 2.0);
    k -= 1;
  }
  const dd s = dd_div(dd_add(q, dd{-1.0, 0.0}), dd_add(q, dd{1.0, 0.0}));
  const dd lm = dd_scale(dd_mul(s, atanh_over_s(s)), 2.0);
  return dd_add(dd_mul(dd{LN2_H, LN2_L}, dd{(double)k, 0.0}), lm);
}

TDC_HD dd log2_ratio(double x1, double x2) { return dd_mul(log_ratio(x1, x2), dd{LOG2E_H, LOG2E_L}); }

// 2^(L.h + L.l) as a double-double: 2^k · e^r with r = (L − k) ln 2 scaled by 2⁻⁸, the Taylor
// series of e^r − 1 to degree 10, then eight squarings of 1 + x
TDC_HD dd exp2_dd(dd L) {
  if (L.h > 1023.0) return dd{INFINITY, 0.0};
  if (L.h < -1074.0) return dd{0.0, 0.0};
  const double k = rint(L.h);
  const dd f = dd_add(L, dd{-k, 0.0});
  const dd r = dd_scale(dd_mul(f, dd{LN2_H, LN2_L}), 1.0 / 256.0);
  dd x = twodiv(1.0, 0.0, 3628800.0, 0.0);  // 1/10!
  double fact = 3628800.0;
  for (int j = 9; j >= 1; --j) {
    fact /= (double)(j + 1);  // j!
    x = dd_add(dd_mul(x, r), twodiv(1.0, 0.0, fact, 0.0));
  }
  x = dd_mul(x, r);  // e^r − 1
  for (int i = 0; i < 8; ++i) x = dd_add(dd_scale(x, 2.0), dd_mul(x, x));
  const dd y = dd_add(dd{1.0, 0.0}, x);
  const int ki = (int)k;
  return dd{ldexp(y.h, ki), ldexp(y.l, ki)};
}

// ------------------------------------------------------------------------------------------
// 3-vectors
// ------------------------------------------------------------------------------------------

TDC_HD double dot3(const double* x, const double* y) { return fma(x[2], y[2], fma(x[1], y[1], x[0] * y[0])); }
TDC_HD double cdot3(const double* x, const double* y) {  // compensated (utils.jl:55-68)
  double s = 0.0, c = 0.0;
  for (int i = 0; i < 3; ++i) {
    double pr, e, e2;
    two_prod(x[i], y[i], pr, e);
    two_sum(s, pr, s, e2);
    c += e + e2;
  }
  return s + c;
}
TDC_HD double norm3(const double* x) { return sqrt(dot3(x, x)); }
TDC_HD void cross3(double* z, const double* x, const double* y) {
  z[0] = fma(x[1], y[2], -x[2] * y[1]);
  z[1] = fma(x[2], y[0], -x[0] * y[2]);
  z[2] = fma(x[0], y[1], -x[1] * y[0]);
}

// f''(p) v = Pᵀ L Lᵀ P v with P v = (v₃, v₁, v₂) (tdc.jl:76-94)
TDC_HD void hessmul(double* out, const double* L, const double* v) {
  const double L11 = L[0], L21 = L[1], L31 = L[2], L22 = L[4], L32 = L[5], L33 = L[8];
  const double r1 = L11 * v[2] + L21 * v[0] + L31 * v[1];
  const double r2 = L22 * v[0] + L32 * v[1];
  const double r3 = L33 * v[1];
  const double t1 = L11 * r1, t2 = L21 * r1 + L22 * r2, t3 = L31 * r1 + L32 * r2 + L33 * r3;
  out[0] = t2;
  out[1] = t3;
  out[2] = t1;
}

// b ← f''(p)⁻¹ b through the pivoted factor (tdc.jl:51-67)
TDC_HD void hessldiv(const double* L, double* b) {
  const double L11 = L[0], L21 = L[1], L31 = L[2], L22 = L[4], L32 = L[5], L33 = L[8];
  const double y1 = b[2] / L11;
  const double y2 = (b[0] - L21 * y1) / L22;
  const double y3 = (b[1] - L31 * y1 - L32 * y2) / L33;
  const double z3 = y3 / L33;
  const double z2 = (y2 - L32 * z3) / L22;
  const double z1 = (y1 - L21 * z2 - L31 * z3) / L11;
  b[0] = z2;
  b[1] = z3;
  b[2] = z1;
}

// H += μ f''(p) = μ Pᵀ L Lᵀ P (tdc.jl:104-119); H column-major 3×3
TDC_HD void gram(double* H, const double* L, double mu) {
  const double L11 = L[0], L21 = L[1], L31 = L[2], L22 = L[4], L32 = L[5], L33 = L[8];
  H[8] += mu * L11 * L11;
  H[0] += mu * (L21 * L21 + L22 * L22);
  H[4] += mu * (L31 * L31 + L32 * L32 + L33 * L33);
  const double h13 = mu * L21 * L11, h23 = mu * L31 * L11, h12 = mu * (L31 * L21 + L32 * L22);
  H[6] += h13;
  H[2] += h13;
  H[7] += h23;
  H[5] += h23;
  H[3] += h12;
  H[1] += h12;
}

// ------------------------------------------------------------------------------------------
// exponential cone {(x₁, x₂, x₃): x₁, x₂ > 0, x₂ log(x₁/x₂) ≥ x₃} (exp.jl)
// barrier f(x) = −log ψ − log x₁ − log x₂, ψ = x₂ log(x₁/x₂) − x₃
// ------------------------------------------------------------------------------------------

constexpr double EXP_E1 = 1.2909282315382298, EXP_E2 = 0.8051015526498357, EXP_E3 = -0.8278379086082098;

// ψ, 1/ψ, log(x₁/x₂), log(x₁/x₂) − 1 through a double-double chain (exp.jl:43-61)
TDC_HD void exp_det(const double* x, double& psi, double& ipsi, double& lr, double& lrm1) {
  const dd r = log_ratio(x[0], x[1]);
  double mh, me;
  two_sum(r.h, -1.0, mh, me);
  lrm1 = mh + (me + r.l);
  double ph, pe;
  two_prod(r.h, x[1], ph, pe);
  pe = fma(r.l, x[1], pe);
  double sh, se;
  two_sum(ph, -x[2], sh, se);
  double qh, ql;
  two_sum(sh, se + pe, qh, ql);
  const dd inv = twodiv(1.0, 0.0, qh, ql);
  psi = qh + ql;
  ipsi = inv.h + inv.l;
  lr = r.h + r.l;
}

TDC_HD void exp_grad(double* g, const double* x, double ipsi, double lrm1) {
  g[0] = -(x[1] * ipsi + 1.0) / x[0];
  g[1] = -lrm1 * ipsi - 1.0 / x[1];
  g[2] = ipsi;
}

TDC_HD bool exp_fact(double* L, const double* x, double psi, double ipsi, double lrm1) {
  const double x1 = x[0], x2 = x[1];
  const double r1 = 1.0 + x2 * ipsi, r2 = (psi + 2.0 * x2) / (psi + x2);
  const bool ok = psi > 0 && r1 > 0 && r2 > 0 && isfinite(r1) && isfinite(r2);
  if (ok) {
    const double s1 = sqrt(r1);
    L[0] = ipsi;
    L[1] = -x2 * ipsi / x1;
    L[2] = -lrm1 * ipsi;
    L[4] = s1 / x1;
    L[5] = -ipsi / s1;
    L[8] = sqrt(r2) / x2;
    L[3] = L[6] = L[7] = 0.0;
  }
  return ok;
}

// f'''(x)[u] as a symmetric 3×3 matrix (exp.jl:105-141)
TDC_HD void exp_third(double* D, const double* x, const double* u, double ipsi, double lrm1) {
  const double x1 = x[0], x2 = x[1], u1 = u[0], u2 = u[1], u3 = u[2];
  const double g1 = x2 / x1, g2 = lrm1;
  const double gu = g1 * u1 + g2 * u2 - u3;
  const double H11 = -x2 / (x1 * x1), H21 = 1.0 / x1, H22 = -1.0 / x2;
  const double Hu1 = H11 * u1 + H21 * u2, Hu2 = H21 * u1 + H22 * u2;
  const double i2 = ipsi * ipsi, i3 = i2 * ipsi;
  const double al = -2.0 * gu * i3;
  const double x1s = x1 * x1, x1c = x1s * x1, x2s = x2 * x2;
  D[0] = al * g1 * g1 + (2.0 * g1 * Hu1 + H11 * gu) * i2 - (2.0 * x2 * u1 / x1c - u2 / x1s) * ipsi - 2.0 * u1 / x1c;
  D[1] = al * g1 * g2 + (g1 * Hu2 + g2 * Hu1 + H21 * gu) * i2 + u1 / x1s * ipsi;
  D[2] = -al * g1 - Hu1 * i2;
  D[4] = al * g2 * g2 + (2.0 * g2 * Hu2 + H22 * gu) * i2 - u2 / x2s * ipsi - 2.0 * u2 / (x2s * x2);
  D[5] = -al * g2 - Hu2 * i2;
  D[8] = al;
  D[3] = D[1];
  D[6] = D[2];
  D[7] = D[5];
}

// the shadow primal p* with −f'(p*) = d, by Halley on the scale-free normal form
// v + log1p(v) = R in v = 1/(p₂*(−d₃)); the root lies in [R/2, R] (exp.jl:199-364).
// Returns the next seed (p₂*) and writes h_dual(0) = 1 − log(−d₃/d₁) + d₂/(−d₃).
TDC_HD double exp_dualgrad(double* sp, double seed, const double* d, double& h0) {
  const double d1 = d[0], d2 = d[1], d3 = d[2];
  if (d3 == 0.0) {
    const double p2 = 1.0 / d2, s = d2 / d1;
    sp[0] = s * p2;
    sp[1] = p2;
    sp[2] = -INFINITY;
    h0 = (d1 >= 0 && d2 >= 0) ? INFINITY : -INFINITY;
    return p2;
  }
  const double md3 = -d3;
  const double rat = md3 / d1;
  const double lr = (rat >= TINY && rat < INFINITY) ? log(rat) : log(md3) - log(d1);
  const double w = d2 / md3;
  double R = w - (lr - 1.0);
  if (!(R > 0)) R = TINY;
  double lo = R / 2, hi = R;
  double v = 1.0 / (fmax(seed, TINY) * md3);
  if (!(lo < v && v < hi)) {
    v = R <= 2 ? R / 2 + R * R / 16 : R - log(R);
    if (!(lo < v && v < hi)) v = 0.7 * R;
  }
  for (int it = 0; it < 60; ++it) {
    const double fv = v + log1p(v) - R;
    const double fp = 1.0 + 1.0 / (1.0 + v);
    if (fv < 0) lo = v;
    else if (fv > 0) hi = v;
    else break;
    const double fpp = -(fp - 1.0) * (fp - 1.0);
    double vn = v - 2.0 * fv * fp / (2.0 * fp * fp - fv * fpp);
    if (!(lo < vn && vn < hi)) {
      if (vn == lo || vn == hi) {
        v = vn;
        break;
      }
      vn = (lo + hi) / 2;
    }
    const bool conv = fabs(vn - v) < 4.0 * EPS * fabs(vn);
    v = vn;
    if (conv) break;
  }
  const double p2 = 1.0 / (md3 * v);
  double s = md3 * (v + 1.0) / d1;
  if (!(0.0 < s && s < INFINITY)) s = exp(lr + log1p(v));
  sp[0] = s * p2;
  sp[1] = p2;
  sp[2] = p2 * log(s) + 1.0 / d3;
  h0 = 1.0 - lr + w;
  return p2;
}

// membership h = log(x₁/x₂) − x₃/x₂ along x = p + τΔp and its first two derivatives
// (exp.jl:421-459)
TDC_HD void exp_jet_prim(double tau, const double* p, const double* dp, double& h, double& hp, double& hpp) {
  const double x1 = fma(tau, dp[0], p[0]), x2 = fma(tau, dp[1], p[1]), x3 = fma(tau, dp[2], p[2]);
  if (x1 > 0 && x2 > 0) {
    const double q = x1 / x2;
    const double s = (q >= TINY && q < INFINITY) ? log(q) : log(x1) - log(x2);
    const double w = x3 / x2;
    h = s - w;
    const double wp = (dp[2] - w * dp[1]) / x2;
    hp = dp[0] / x1 - dp[1] / x2 - wp;
    hpp = -(dp[0] / x1) * (dp[0] / x1) + (dp[1] / x2) * (dp[1] / x2) + 2.0 * wp * dp[1] / x2;
    return;
  }
  hp = hpp = NAN;
  h = (x1 >= 0 && x2 == 0.0) ? (x3 <= 0 ? INFINITY : -INFINITY) : -INFINITY;
}

// dual membership h = 1 + log z₁ − z₂/z₃ − log(−z₃) along z = d + τΔd (exp.jl:468-495)
TDC_HD void exp_jet_dual(double tau, const double* d, const double* dd_, double& h, double& hp, double& hpp) {
  const double z1 = fma(tau, dd_[0], d[0]), z2 = fma(tau, dd_[1], d[1]), z3 = fma(tau, dd_[2], d[2]);
  if (z1 > 0 && z3 < 0) {
    const double r = z2 / z3;
    h = 1.0 + log(z1) - r - log(-z3);
    const double rp = (dd_[1] - r * dd_[2]) / z3;
    hp = dd_[0] / z1 - rp - dd_[2] / z3;
    hpp = -(dd_[0] / z1) * (dd_[0] / z1) + 2.0 * rp * dd_[2] / z3 + (dd_[2] / z3) * (dd_[2] / z3);
    return;
  }
  hp = hpp = NAN;
  h = (z1 >= 0 && z2 >= 0 && z3 == 0.0) ? INFINITY : -INFINITY;
}

// ------------------------------------------------------------------------------------------
// power cone {(x₁, x₂, x₃): x₁, x₂ ≥ 0, x₁^α x₂^(1−α) ≥ |x₃|} (pow.jl)
// barrier f(x) = −log(x₁^{2α} x₂^{2−2α} − x₃²) − (1−α) log x₁ − α log x₂
// ------------------------------------------------------------------------------------------

// P = x₁^{2α} x₂^{2−2α}, ρ = P/det, 1/det with det = P − x₃², in double-double (pow.jl:47-70)
TDC_HD void pow_det(const double* x, double al, double& P, double& rho, double& idet) {
  const double a = 2.0 * al;
  const dd r = log2_ratio(x[0], x[1]);
  const dd g = log2_ratio(x[1], 1.0);
  const dd L2 = dd_add(dd_mul(dd{a, 0.0}, r), dd_scale(g, 2.0));
  const dd Pd = exp2_dd(L2);
  double qh, ql;
  two_prod(x[2], x[2], qh, ql);
  double fh, fe;
  two_sum(Pd.h, -qh, fh, fe);
  double dh, dl;
  two_sum(fh, fe + (Pd.l - ql), dh, dl);
  const dd inv = twodiv(1.0, 0.0, dh, dl);
  double rh, re;
  two_prod(Pd.h, inv.h, rh, re);
  re = fma(Pd.h, inv.l, fma(Pd.l, inv.h, re));
  double rl;
  two_sum(rh, re, rh, rl);
  P = Pd.h + Pd.l;
  rho = rh + rl;
  idet = inv.h + inv.l;
}

TDC_HD void pow_grad(double* g, const double* x, double al, double rho, double idet) {
  const double a = 2.0 * al, b = 2.0 - 2.0 * al;
  g[0] = -(a * rho + 1.0 - al) / x[0];
  g[1] = -(b * rho + al) / x[1];
  g[2] = 2.0 * x[2] * idet;
}

TDC_HD bool pow_fact(double* L, const double* x, double al, double P, double rho, double idet) {
  const double x1 = x[0], x2 = x[1], x3 = x[2];
  const double a = 2.0 * al, b = 2.0 - 2.0 * al;
  const double w = rho * idet, s = P + x3 * x3;
  const double l1 = a / x1, l2 = b / x2;
  const double d1 = (2.0 * rho * a + b) / (2.0 * x1 * x1);
  const double d2 = (2.0 * rho * b + a) / (2.0 * x2 * x2);
  const double c = rho * x3 * x3 / s;
  const double r1 = d1 - c * l1 * l1;
  const double r2 = (d1 * d2 - c * (l2 * l2 * d1 + l1 * l1 * d2)) / r1;
  const bool ok = r1 > 0 && r2 > 0 && isfinite(r2);
  if (ok) {
    L[0] = sqrt(2.0 * s) * idet;
    L[1] = -2.0 * l1 * x3 * w / L[0];
    L[2] = -2.0 * l2 * x3 * w / L[0];
    L[4] = sqrt(r1);
    L[8] = sqrt(r2);
    L[5] = -c * l1 * l2 / L[4];
    L[3] = L[6] = L[7] = 0.0;
  }
  return ok;
}

// f'''(x)[u] (pow.jl:139-191)
TDC_HD void pow_third(double* D, const double* x, const double* u, double al, double P, double rho, double idet) {
  const double x1 = x[0], x2 = x[1], x3 = x[2], u1 = u[0], u2 = u[1], u3 = u[2];
  const double a = 2.0 * al, b = 2.0 - 2.0 * al;
  const double l1 = a / x1, l2 = b / x2, m1 = u1 / x1, m2 = u2 / x2;
  const double f1 = P * l1, f2 = P * l2, f3 = -2.0 * x3;
  const double fdot = f1 * u1 + f2 * u2 + f3 * u3;
  const double f11 = P * l1 * (l1 - 1.0 / x1), f22 = P * l2 * (l2 - 1.0 / x2), f12 = P * l1 * l2;
  const double fd1 = f11 * u1 + f12 * u2, fd2 = f12 * u1 + f22 * u2, fd3 = -2.0 * u3;
  const double i2 = idet * idet, i3 = i2 * idet;
  const double c2 = fdot * i2, c3 = 2.0 * fdot * i3;
  const double K = rho * a * b * (a - 1.0) * (m2 - m1);
  D[0] = 2.0 * fd1 * f1 * i2 - c3 * f1 * f1 - K / (x1 * x1) + c2 * f11 - b * m1 / (x1 * x1);
  D[4] = 2.0 * fd2 * f2 * i2 - c3 * f2 * f2 - K / (x2 * x2) + c2 * f22 - a * m2 / (x2 * x2);
  D[1] = (fd2 * f1 + f2 * fd1) * i2 - c3 * f2 * f1 + K / (x1 * x2) + c2 * f12;
  D[8] = 2.0 * fd3 * f3 * i2 - c3 * f3 * f3 - 2.0 * c2;
  D[2] = (fd3 * f1 + f3 * fd1) * i2 - c3 * f3 * f1;
  D[5] = (fd3 * f2 + f3 * fd2) * i2 - c3 * f3 * f2;
  D[3] = D[1];
  D[6] = D[2];
  D[7] = D[5];
}

// the shadow primal by the log residual of the normal form
// 1 − u = W₀ (1 + k₁u)^a (1 + k₂u)^b in u = 1/ρ ∈ (0, 1], root in [(1−W₀)/2, 1−W₀]
// (pow.jl:262-447). Returns the next seed ρ and writes h_dual(0) = −log(W₀)/2.
TDC_HD double pow_dualgrad(double* sp, double seed, const double* d, double al, double& h0) {
  const double d1 = d[0], d2 = d[1], d3 = d[2];
  const double a = 2.0 * al, b = 2.0 - 2.0 * al;
  if (d3 == 0.0) {
    sp[0] = (a + 1.0 - al) / d1;
    sp[1] = (b + al) / d2;
    sp[2] = 0.0;
    h0 = INFINITY;
    return 1.0;
  }
  const double k1 = (1.0 - al) / a, k2 = al / b;
  const double t1 = a / d1, t2 = b / d2;
  const double z1 = (d3 / 2.0) * t2, z1s = z1 * z1, z2 = t1 / t2, z2a = pow(z2, a);
  double W0 = z1s * z2a, lW0;
  const bool normal = z1s >= TINY && z1s < INFINITY && z2 >= TINY && z2 < INFINITY && z2a >= TINY &&
                      z2a < INFINITY && W0 >= TINY && W0 < INFINITY;
  if (!normal) {
    lW0 = 2.0 * log(fabs(d3)) - log(4.0) + a * (log(a) - log(d1)) + b * (log(b) - log(d2));
    W0 = exp(lW0);
  } else {
    lW0 = log(W0);
  }
  double f0 = 1.0 - W0;
  if (!(f0 > 0)) f0 = TINY;
  double lo = f0 / 2, hi = f0;
  double u = 1.0 / fmax(seed, 1.0);
  if (!(lo < u && u < hi)) {
    const double Fpp0 = 1.0 - (k1 * k1 * a + k2 * k2 * b);
    u = -4.0 * lW0 / (8.0 - lW0 * Fpp0);
    if (!(lo < u && u < hi)) u = 0.7 * f0;
  }
  for (int it = 0; it < 60; ++it) {
    const double r1 = 1.0 + k1 * u, r2 = 1.0 + k2 * u;
    const double Fv = lW0 + a * log1p(k1 * u) + b * log1p(k2 * u) - log1p(-u);
    const double sg = k1 * a / r1 + k2 * b / r2;
    const double Fp = sg + 1.0 / (1.0 - u);
    if (Fv < 0) lo = u;
    else if (Fv > 0) hi = u;
    else break;
    const double sgp = -(k1 * k1 * a / (r1 * r1) + k2 * k2 * b / (r2 * r2));
    const double Fpp = sgp + 1.0 / ((1.0 - u) * (1.0 - u));
    const double halley = u - 2.0 * Fv * Fp / (2.0 * Fp * Fp - Fv * Fpp);
    const double pole = 1.0 - (1.0 - u) * exp(Fv);
    double un, alt;
    if (1.0 / (1.0 - u) > 4.0 * sg) {
      un = pole;
      alt = halley;
    } else {
      un = halley;
      alt = pole;
    }
    if (!(lo < un && un < hi)) {
      if (un == lo || un == hi) {
        u = un;
        break;
      }
      un = (lo < alt && alt < hi) ? alt : (lo + hi) / 2;
    }
    const bool conv = fabs(un - u) < 4.0 * EPS * (1.0 + fabs(un));
    u = un;
    if (conv) break;
  }
  const double rho = 1.0 / u;
  sp[0] = (a * rho + 1.0 - al) / d1;
  sp[1] = (b * rho + al) / d2;
  sp[2] = -2.0 * (rho - 1.0) / d3;
  h0 = -lW0 / 2.0;
  return rho;
}

// h = α log x₁ + (1−α) log x₂ − log|x₃| along x = p + τΔp (pow.jl:495-524)
TDC_HD void pow_jet_prim(double tau, const double* p, const double* dp, double al, double& h, double& hp,
                         double& hpp) {
  const double be = 1.0 - al;
  const double x1 = fma(tau, dp[0], p[0]), x2 = fma(tau, dp[1], p[1]), x3 = fma(tau, dp[2], p[2]);
  hp = hpp = NAN;
  if (x1 > 0 && x2 > 0) {
    if (x3 == 0.0) {
      h = INFINITY;
      return;
    }
    h = al * log(x1) + be * log(x2) - log(fabs(x3));
    hp = al * dp[0] / x1 + be * dp[1] / x2 - dp[2] / x3;
    hpp = -al * (dp[0] / x1) * (dp[0] / x1) - be * (dp[1] / x2) * (dp[1] / x2) + (dp[2] / x3) * (dp[2] / x3);
    return;
  }
  if ((x1 == 0.0 && x2 >= 0) || (x1 >= 0 && x2 == 0.0)) h = x3 == 0.0 ? 0.0 : -INFINITY;
  else h = -INFINITY;
}

// h = α log(s₁/α) + (1−α) log(s₂/(1−α)) − log|s₃| along s = d + τΔd (pow.jl:533-560)
TDC_HD void pow_jet_dual(double tau, const double* d, const double* dd_, double al, double& h, double& hp,
                         double& hpp) {
  const double be = 1.0 - al;
  const double s1 = fma(tau, dd_[0], d[0]), s2 = fma(tau, dd_[1], d[1]), s3 = fma(tau, dd_[2], d[2]);
  hp = hpp = NAN;
  if (s1 > 0 && s2 > 0) {
    if (s3 == 0.0) {
      h = INFINITY;
      return;
    }
    h = al * log(s1 / al) + be * log(s2 / be) - log(fabs(s3));
    hp = al * dd_[0] / s1 + be * dd_[1] / s2 - dd_[2] / s3;
    hpp = -al * (dd_[0] / s1) * (dd_[0] / s1) - be * (dd_[1] / s2) * (dd_[1] / s2) + (dd_[2] / s3) * (dd_[2] / s3);
    return;
  }
  if ((s1 == 0.0 && s2 >= 0) || (s1 >= 0 && s2 == 0.0)) h = s3 == 0.0 ? 0.0 : -INFINITY;
  else h = -INFINITY;
}

// ------------------------------------------------------------------------------------------
// cone-generic layer (tdc.jl)
// ------------------------------------------------------------------------------------------

TDC_HD void identity(int kind, double al, double* x) {
  if (kind == CONE_EXP) {
    x[0] = EXP_E1;
    x[1] = EXP_E2;
    x[2] = EXP_E3;
  } else {
    x[0] = sqrt(1.0 + al);
    x[1] = sqrt(2.0 - al);
    x[2] = 0.0;
  }
}

TDC_HD double seed0(int kind) { return kind == CONE_EXP ? EXP_E2 : 1.0; }

// determinant chain into the cache and h_prim(0) (tddet!: exp.jl:374-389, pow.jl:456-470)
TDC_HD void det(int kind, double al, const double* p, double* c) {
  if (kind == CONE_EXP) {
    double psi, ipsi, lr, lrm1;
    exp_det(p, psi, ipsi, lr, lrm1);
    c[C_D1] = psi;
    c[C_D2] = ipsi;
    c[C_D3] = lr;
    c[C_D4] = lrm1;
    c[C_H0P] = lr - p[2] / p[1];
  } else {
    double P, rho, idet;
    pow_det(p, al, P, rho, idet);
    c[C_D1] = P;
    c[C_D2] = rho;
    c[C_D3] = idet;
    c[C_H0P] = log(P) / 2.0 - log(fabs(p[2]));
  }
}

TDC_HD void barrgrad(int kind, double al, double* g, const double* p, const double* c) {
  if (kind == CONE_EXP) exp_grad(g, p, c[C_D2], c[C_D4]);
  else pow_grad(g, p, al, c[C_D2], c[C_D3]);
}

TDC_HD bool fact(int kind, double al, double* L, const double* p, const double* c) {
  return kind == CONE_EXP ? exp_fact(L, p, c[C_D1], c[C_D2], c[C_D4]) : pow_fact(L, p, al, c[C_D1], c[C_D2], c[C_D3]);
}

TDC_HD void third(int kind, double al, double* D, const double* p, const double* u, const double* c) {
  if (kind == CONE_EXP) exp_third(D, p, u, c[C_D2], c[C_D4]);
  else pow_third(D, p, u, al, c[C_D1], c[C_D2], c[C_D3]);
}

TDC_HD double dualgrad(int kind, double al, double* sp, double seed, const double* d, double* c) {
  double h0;
  const double next = kind == CONE_EXP ? exp_dualgrad(sp, seed, d, h0) : pow_dualgrad(sp, seed, d, al, h0);
  c[C_H0D] = h0;
  return next;
}

// t of the rank-one term t z zᵀ (tdbfgs, tdc.jl:278-354)
TDC_HD double bfgs(const double* L, const double* sp, const double* sd, const double* z, const double* p, double mu,
                   double mut) {
  double w[3], Hw[3], Hz[3];
  for (int i = 0; i < 3; ++i) w[i] = sp[i] - mut * p[i];
  hessmul(Hw, L, w);
  const double wHw = dot3(w, Hw);
  if (!(wHw > 0)) return 0.0;
  hessmul(Hz, L, z);
  const double fzz = dot3(z, Hz), sdz = dot3(sd, z), wHz = dot3(Hw, z);
  double s1, e1, s2, e2;
  two_sum(fzz, -sdz * sdz / 3.0, s1, e1);
  two_sum(s1, -wHz * wHz / wHw, s2, e2);
  return mu * (s2 + (e1 + e2));
}

// Tunçel scaling block H (column-major 3×3, overwritten) and the cache (tdscale!, tdc.jl:185-270).
// Returns the flag; spsd ← ⟨p*, d*⟩.
TDC_HD bool scale(int kind, double al, const double* p, const double* d, double* c, double* H, double& spsd) {
  for (int i = 0; i < 9; ++i) H[i] = 0.0;
  spsd = 0.0;
  det(kind, al, p, c);
  double* sd = c + C_SD;
  barrgrad(kind, al, sd, p, c);
  for (int i = 0; i < 3; ++i) sd[i] = -sd[i];
  double* L = c + C_L;
  const bool ok = fact(kind, al, L, p, c);
  if (!ok) return false;
  double sp[3];
  c[C_SEED] = dualgrad(kind, al, sp, c[C_SEED], d, c);
  const double mu = dot3(p, d) / 3.0;
  spsd = cdot3(sp, sd);
  const double mut = spsd / 3.0;
  double z[3];
  cross3(z, p, sp);
  const double nz = norm3(z), np = norm3(p), nsp = norm3(sp);
  if (nz < EPS * (np * nsp + EPS)) {
    gram(H, L, mu);
    return true;
  }
  for (int i = 0; i < 3; ++i) z[i] /= nz;
  const double t = bfgs(L, sp, sd, z, p, mu, mut);
  if (t <= 0 || !isfinite(t)) {
    gram(H, L, mu);
    return true;
  }
  double dp[3], dd_[3];
  for (int i = 0; i < 3; ++i) {
    dp[i] = p[i] - mu * sp[i];
    dd_[i] = d[i] - mu * sd[i];
  }
  const double pd = 3.0 * mu, dpd = cdot3(dp, dd_);
  const double ipd = 1.0 / pd, idpd = 1.0 / dpd;
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) H[i + 3 * j] += ipd * d[i] * d[j] + idpd * dd_[i] * dd_[j] + t * z[i] * z[j];
  return true;
}

// predictor right-hand side σμ d* − d (tdcorr0!, tdc.jl:414-419)
TDC_HD void corr0(double* r, const double* d, const double* c, double smu) {
  for (int i = 0; i < 3; ++i) r[i] = smu * c[C_SD + i] - d[i];
}

// Mehrotra corrector −d + σμ d* − η, η = −½ f'''(p)[Δp, f''(p)⁻¹Δd] (tdcorr!, tdc.jl:371-408)
TDC_HD void corr(int kind, double al, double* r, const double* p, const double* d, const double* dp,
                 const double* dd_, const double* c, double smu) {
  double v[3] = {dd_[0], dd_[1], dd_[2]};
  hessldiv(c + C_L, v);
  double D[9];
  third(kind, al, D, p, dp, c);
  for (int i = 0; i < 3; ++i) {
    const double eta = -0.5 * (D[i] * v[0] + D[i + 3] * v[1] + D[i + 6] * v[2]);
    r[i] = smu * c[C_SD + i] - d[i] - eta;
  }
}

// largest affine step keeping the strictly signed coordinates positive (exp.jl:151-193,
// pow.jl:200-240)
TDC_HD double bound(const double* x, const double* dx, int i0, int i1, bool neg1) {
  double hi = 1.0;
  if (dx[i0] < 0) hi = fmin(hi, -x[i0] / dx[i0]);
  if (neg1 ? dx[i1] > 0 : dx[i1] < 0) hi = fmin(hi, -x[i1] / dx[i1]);
  return hi;
}

// certified step τ ∈ [0, hi] with h(τ) ≥ band, within NFK band widths of the last root of
// the log-form membership h (nflog3, tdc/utils.jl:24-200): chord / logarithmic / pole models
// selected against h′, matched one-point boundary models, Halley / Newton fallbacks, and a
// forced bisection when a round does not halve the bracket
template <class Jet>
TDC_HD double nflog3(Jet jet, double hi, double glo, double gplo) {
  const double tb = hi, nb = NB_BAND;
  double lo = 0.0;
  if (!(glo >= nb)) return lo;
  double ghi, gphi, gpphi;
  jet(hi, ghi, gphi, gpphi);
  if (ghi >= nb) return hi;
  double whi = (isfinite(gphi) && gphi != 0.0) ? nb / fabs(gphi) : 0.0;
  auto take = [&](double m) -> bool {  // evaluate at m, shrink the bracket; true when resolved
    double g, gp, gpp;
    jet(m, g, gp, gpp);
    if (g >= nb) {
      lo = m;
      glo = g;
      gplo = gp;
    } else {
      hi = m;
      ghi = g;
      gphi = gp;
      gpphi = gpp;
      whi = (isfinite(gp) && gp != 0.0) ? nb / fabs(gp) : 0.0;
    }
    return hi - lo < fmax(NFK * whi, EPS * (1.0 + hi));
  };
  for (int it = 0; it < 53; ++it) {
    const double wid = hi - lo;
    double m;
    if (isfinite(ghi) && ghi >= -nb) {
      m = hi - 3.0 * fmax(whi, EPS * (1.0 + hi)) / 2.0;
    } else if (isfinite(ghi) && isfinite(glo)) {
      m = lo + glo * wid / (glo - ghi);
      double best = fabs((ghi - glo) / wid - gplo);
      const double dlo = tb - lo, dhi = tb - hi;
      if (dhi > 0) {
        const double ulo = log(dlo), uhi = log(dhi);
        const double cc = (glo - ghi) / (ulo - uhi);
        double e = fabs(-cc / dlo - gplo);
        if (e < best) {
          best = e;
          const double m2 = tb - exp(ulo - glo * (ulo - uhi) / (glo - ghi));
          if (lo < m2 && m2 < hi) m = m2;
        }
        const double bb = (glo - ghi) / (1.0 / dhi - 1.0 / dlo);
        const double aa = glo + bb / dlo;
        e = fabs(-bb / (dlo * dlo) - gplo);
        if (e < best && aa != 0.0) {
          const double m2 = tb - bb / aa;
          if (lo < m2 && m2 < hi) m = m2;
        }
      }
    } else {
      m = lo + wid / 2;
    }
    if (!(lo < m && m < hi)) m = lo + wid / 2;
    if (take(m)) break;
    const double dh = tb - hi;
    m = NAN;
    if (isfinite(gpphi) && isfinite(gphi) && gphi != 0.0 && dh > 0) {
      const double rr = gpphi * dh / gphi;
      if (0.5 < rr && rr < 1.5) {
        m = tb - dh * exp(ghi / (gphi * dh));
      } else if (1.5 <= rr && rr < 3.0) {
        const double den = ghi - gphi * dh;
        if (den != 0.0) m = tb + gphi * dh * dh / den;
      }
    }
    if (!(lo < m && m < hi)) {
      m = hi - 2.0 * ghi * gphi / (2.0 * gphi * gphi - ghi * gpphi);
      if (!(lo < m && m < hi)) m = hi - ghi / gphi;
    }
    if (!(lo < m && m < hi)) m = lo + (hi - lo) / 2;
    if (take(m)) break;
    if (2.0 * (hi - lo) > wid) {
      if (take(lo + (hi - lo) / 2)) break;
    }
  }
  return lo;
}

// (τp, τd) before step_frac (tdmaxsteps, tdc.jl:437-458)
TDC_HD void maxsteps(int kind, double al, const double* p, const double* dp, const double* d, const double* dd_,
                     const double* c, double& tp, double& td) {
  if (kind == CONE_EXP) {
    const double hip = bound(p, dp, 0, 1, false), hid = bound(d, dd_, 0, 2, true);
    const double w = p[2] / p[1];
    const double hp0 = dp[0] / p[0] - dp[1] / p[1] - (dp[2] - w * dp[1]) / p[1];
    const double r = d[1] / d[2];
    const double hd0 = dd_[0] / d[0] - (dd_[1] - r * dd_[2]) / d[2] - dd_[2] / d[2];
    tp = nflog3([&](double t, double& h, double& g, double& gg) { exp_jet_prim(t, p, dp, h, g, gg); }, hip,
                c[C_H0P], hp0);
    td = nflog3([&](double t, double& h, double& g, double& gg) { exp_jet_dual(t, d, dd_, h, g, gg); }, hid,
                c[C_H0D], hd0);
  } else {
    const double hip = bound(p, dp, 0, 1, false), hid = bound(d, dd_, 0, 1, false);
    const double hp0 = al * dp[0] / p[0] + (1.0 - al) * dp[1] / p[1] - dp[2] / p[2];
    const double hd0 = al * dd_[0] / d[0] + (1.0 - al) * dd_[1] / d[1] - dd_[2] / d[2];
    tp = nflog3([&](double t, double& h, double& g, double& gg) { pow_jet_prim(t, p, dp, al, h, g, gg); }, hip,
                c[C_H0P], hp0);
    td = nflog3([&](double t, double& h, double& g, double& gg) { pow_jet_dual(t, d, dd_, al, h, g, gg); }, hid,
                c[C_H0D], hd0);
  }
}

// s = −f'(x) (dualshadow!) and x = (−f')⁻¹(s) (primalshadow!) (tdc.jl:464-478); c: a
// scratch cache
TDC_HD void dual_shadow(int kind, double al, double* s, const double* x, double* c) {
  det(kind, al, x, c);
  barrgrad(kind, al, s, x, c);
  for (int i = 0; i < 3; ++i) s[i] = -s[i];
}
TDC_HD void primal_shadow(int kind, double al, double* x, const double* s, double* c) {
  (void)dualgrad(kind, al, x, 0.0, s, c);
}

#undef TDC_HD

}  // namespace tdc
}  // namespace mb
