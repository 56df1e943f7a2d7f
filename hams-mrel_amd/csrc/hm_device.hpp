// hm_device.hpp -- FP64 device math of the boundary-integral hot path (gfx950).
//
// Special functions (src/FinGrnExtSubs.f90), the finite-depth wave Green
// function (src/FinGreen3D.f90), the infinite-depth Havelock approximation
// (src/InfGreen_Appr.f90) and the Rankine panel integrals (src/SingularIntgr.f90),
// written for one GPU thread per (field point, source panel) pair:
//   * complex values are carried as separate real/imaginary doubles; in every
//     regime only the propagating-mode terms have an imaginary part;
//   * each R/h regime has its own entry point so that a launch can be restricted
//     to pairs of one regime (and one series length), keeping waves convergent;
//   * the adaptive Gauss-Kronrod quadrature of the Linton regime is an explicit
//     interval stack instead of recursion;
//   * the Wynn-epsilon tables live in per-lane arrays sized for the longest
//     series of each regime (51 terms for R/h >= 0.05, 100 below).
// REAL(4) literals of the reference (thresholds, term-count fits, one JY01B
// coefficient) are reproduced as float-rounded doubles, F32(...), because they
// select branches and series lengths.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define HM_DEV __device__ __host__ inline
#define F32(x) ((double)(float)(x))

namespace hm {

constexpr double kPi = 3.141592653589793;
constexpr int kEigeneMaxTerms = 51;   // INT(-88.89*0.05+54.45) = 50 partial sums + 1
constexpr int kPidcockMaxTerms = 100; // INT(-1010.10*0.0005+100.50) = 99 partial sums + 1

struct cplx { double re, im; };

// ---------------------------------------------------------------------------
// Special functions
// ---------------------------------------------------------------------------

// erf, series for |x|<3.5 else asymptotic (ERROR, FinGrnExtSubs.f90:62)
HM_DEV double sf_erf(double x) {
  const double x2 = x * x;
  if (fabs(x) < 3.5) {
    double er = 1.0, r = 1.0;
    for (int k = 1; k <= 50; ++k) {
      r = r * x2 / (k + 0.5);
      er = er + r;
      if (fabs(r) <= fabs(er) * 1.0e-15) break;
    }
    return 2.0 / sqrt(kPi) * x * exp(-x2) * er;
  }
  double er = 1.0, r = 1.0;
  for (int k = 1; k <= 12; ++k) {
    r = -r * (k - 0.5) / x2;
    er = er + r;
  }
  double e = 1.0 - exp(-x2) / (fabs(x) * sqrt(kPi)) * er;
  return x < 0.0 ? -e : e;
}
HM_DEV double sf_erfc(double x) { return 1.0 - sf_erf(x); }
HM_DEV double sf_derfc(double x) { return -2.0 / sqrt(kPi) * exp(-(x * x)); }

// E1(x) rational fits (E1XA :198)
HM_DEV double sf_e1(double x) {
  if (x == 0.0) return 1.0e300;
  if (x <= 1.0)
    return -log(x) + ((((1.07857e-3 * x - 9.76004e-3) * x + 5.519968e-2) * x - 0.24991055) * x +
                      0.99999193) * x - 0.57721566;
  double es1 = (((x + 8.5733287401) * x + 18.059016973) * x + 8.6347608925) * x + 0.2677737343;
  double es2 = (((x + 9.5733223454) * x + 25.6329561486) * x + 21.0996530827) * x + 3.9584969228;
  return exp(-x) / x * es1 / es2;
}

// En(x) by upward recurrence (ENXA :135)
HM_DEV double sf_en(int n, double x) {
  double e1 = sf_e1(x);
  if (n == 0) return exp(-x) / x;
  if (n == 1) return e1;
  double en = e1, ex = exp(-x);
  for (int k = 2; k <= n; ++k) {
    en = (ex - x * e1) / (k - 1.0);
    e1 = en;
  }
  return en;
}

// Ei(x) (EIX :254)
HM_DEV double sf_ei(double x) {
  if (x == 0.0) return -1.0e300;
  if (x <= 40.0) {
    double ei = 1.0, r = 1.0;
    for (int k = 1; k <= 100; ++k) {
      r = r * k * x / ((k + 1.0) * (k + 1.0));
      ei = ei + r;
      if (fabs(r / ei) <= 1.0e-15) break;
    }
    return 0.5772156649015328 + log(x) + x * ei;
  }
  double ei = 1.0, r = 1.0;
  for (int k = 1; k <= 20; ++k) {
    r = r * k / x;
    ei = ei + r;
  }
  return exp(x) / x * ei;
}

// Correctly rounded quotients by the integer divisors of the series below,
// without the generic division sequence: with y = RN(1/b) from a table rounded
// at compile time, q = RN(a*y) is within an ulp, the remainder r = a - b*q is
// exact with an FMA, and RN(q + r*y) is the correctly rounded quotient
// (Markstein's correction step).
struct RcpTables { double k[51], kk[51], kk1[51]; };
constexpr RcpTables make_rcp_tables() {
  RcpTables t{};
  for (int i = 1; i <= 50; ++i) {
    t.k[i] = 1.0 / i;
    t.kk[i] = 1.0 / ((double)i * i);
    t.kk1[i] = 1.0 / ((double)i * (i + 1));
  }
  return t;
}
static __constant__ constexpr RcpTables kRcp = make_rcp_tables();

HM_DEV double div_by(double a, double b, double y) {
  const double q = a * y;
  const double r = fma(-b, q, a);
  return fma(r, y, q);
}
// 1/b from the hardware estimate and two Newton steps, then a / b as above
HM_DEV double div_cr(double a, double b) {
#ifdef __HIP_DEVICE_COMPILE__
  double y = __builtin_amdgcn_rcp(b);
  double e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  return div_by(a, b, y);
#else
  return a / b;
#endif
}

// K0(x), K1(x) (IK01A :538). The I0/I1 series feed K1 and the large-x K0.
// Same operations in the same order as the reference, without multiply-add
// contraction (the K0 series cancels by up to e^(2x), so one rounding more or
// less shows at 1e-9); the divisions inside the loops are the correctly rounded
// quotients of div_by and div_cr.
HM_DEV void sf_k01(double x, double& bk0, double& bk1) {
#pragma clang fp contract(off)
  if (x == 0.0) { bk0 = 1.0e300; bk1 = 1.0e300; return; }
  const double x2 = x * x;
  double bi0, bi1;
  if (x <= 18.0) {
    double r = 1.0;
    bi0 = 1.0;
    for (int k = 1; k <= 50; ++k) {
      r = div_by(0.25 * r * x2, (double)(k * k), kRcp.kk[k]);
      bi0 += r;
      if (fabs(div_cr(r, bi0)) < 1.0e-15) break;
    }
    r = 1.0;
    bi1 = 1.0;
    for (int k = 1; k <= 50; ++k) {
      r = div_by(0.25 * r * x2, (double)(k * (k + 1)), kRcp.kk1[k]);
      bi1 += r;
      if (fabs(div_cr(r, bi1)) < 1.0e-15) break;
    }
    bi1 = 0.5 * x * bi1;
  } else {
    const double a[12] = {0.125, 7.03125e-2, 7.32421875e-2, 1.1215209960938e-1,
                          2.2710800170898e-1, 5.7250142097473e-1, 1.7277275025845,
                          6.0740420012735, 2.4380529699556e01, 1.1001714026925e02,
                          5.5133589612202e02, 3.0380905109224e03};
    const double b[12] = {-0.375, -1.171875e-1, -1.025390625e-1, -1.4419555664063e-1,
                          -2.7757644653320e-1, -6.7659258842468e-1, -1.9935317337513,
                          -6.8839142681099, -2.7248827311269e01, -1.2159789187654e02,
                          -6.0384407670507e02, -3.3022722944809e03};
    int k0 = x >= 50.0 ? 7 : (x >= 35.0 ? 9 : 12);
    double ca = exp(x) / sqrt(2.0 * kPi * x), xr = 1.0 / x, p = 1.0;
    bi0 = 1.0;
    bi1 = 1.0;
    for (int k = 0; k < k0; ++k) {
      p *= xr;
      bi0 += a[k] * p;
      bi1 += b[k] * p;
    }
    bi0 *= ca;
    bi1 *= ca;
  }
  double k0v;
  if (x <= 9.0) {
    const double ct = -(log(x / 2.0) + 0.5772156649015329);
    double w0 = 0.0, ww = 0.0, r = 1.0;
    k0v = 0.0;
    for (int k = 1; k <= 50; ++k) {
      w0 += kRcp.k[k];
      r = div_by(0.25 * r, (double)(k * k), kRcp.kk[k]) * x2;
      k0v += r * (w0 + ct);
      if (fabs(div_cr(k0v - ww, k0v)) < 1.0e-15) break;
      ww = k0v;
    }
    k0v += ct;
  } else {
    const double a1[8] = {0.125, 0.2109375, 1.0986328125, 1.1775970458984e01,
                          2.1461706161499e02, 5.9511522710323e03, 2.3347645606175e05,
                          1.2312234987631e07};
    double xr2 = 1.0 / x2, p = 1.0;
    k0v = 1.0;
    for (int k = 0; k < 8; ++k) {
      p *= xr2;
      k0v += a1[k] * p;
    }
    k0v = 0.5 / x * k0v / bi0;
  }
  bk0 = k0v;
  bk1 = (1.0 / x - bi1 * k0v) / bi0;
}

// K0(x), K1(x) for the mixed-precision paths: the same series and expansions as
// sf_k01 (IK01A), with the power series evaluated as fixed-degree polynomials
// in t = x^2 / 4 by Horner's rule instead of term-by-term loops with a
// division and a convergence test per term (A&S 9.6.10, 9.6.11, 9.6.13):
//   I0 = sum t^k / (k!)^2,  I1 = (x / 2) sum t^k / (k! (k+1)!),
//   K0 = -(ln(x/2) + gamma) I0 + sum_{k>=1} H_k t^k / (k!)^2   (x <= 9),
// K0 = (1 / 2x)(1 + sum a_k x^-2k) / I0 (the I0 K0 product expansion) above 9,
// and K1 = (1/x - I1 K0) / I0 from the Wronskian. Degree 28 covers x <= 9
// (t <= 20.25) and degree 42 x <= 18 to below 1e-17 of I0; above 18, I0 and I1
// come from their asymptotic expansions as in sf_k01. The results differ from
// sf_k01 by rounding only (the K0 series cancels near x = 9 in both).
struct BesselSeries { double i0[51], i1[51], k0[51]; };
constexpr BesselSeries make_bessel_series() {
  BesselSeries b{};
  double f2 = 1.0, f1 = 1.0, hk = 0.0;   // (k!)^2, k! (k+1)!, H_k
  b.i0[0] = 1.0;
  b.i1[0] = 1.0;
  b.k0[0] = 0.0;
  for (int k = 1; k <= 50; ++k) {
    f2 *= (double)k * k;
    f1 *= (double)k * (k + 1);
    hk += 1.0 / k;
    b.i0[k] = 1.0 / f2;
    b.i1[k] = 1.0 / f1;
    b.k0[k] = hk / f2;
  }
  return b;
}
static __constant__ constexpr BesselSeries kBes = make_bessel_series();

template <int N>
HM_DEV double horner(const double* c, double t) {
  double s = c[N];
#pragma unroll
  for (int k = N - 1; k >= 0; --k) s = fma(s, t, c[k]);
  return s;
}

HM_DEV void sf_k01_fast(double x, double& bk0, double& bk1) {
  if (x == 0.0) { bk0 = 1.0e300; bk1 = 1.0e300; return; }
  const double t = 0.25 * x * x;
  double bi0, bi1, k0v;
  if (x <= 9.0) {
    bi0 = horner<28>(kBes.i0, t);
    bi1 = 0.5 * x * horner<28>(kBes.i1, t);
    k0v = -(log(0.5 * x) + 0.5772156649015329) * bi0 + horner<28>(kBes.k0, t);
  } else {
    if (x <= 18.0) {
      bi0 = horner<42>(kBes.i0, t);
      bi1 = 0.5 * x * horner<42>(kBes.i1, t);
    } else {
      const double a[12] = {0.125, 7.03125e-2, 7.32421875e-2, 1.1215209960938e-1,
                            2.2710800170898e-1, 5.7250142097473e-1, 1.7277275025845,
                            6.0740420012735, 2.4380529699556e01, 1.1001714026925e02,
                            5.5133589612202e02, 3.0380905109224e03};
      const double b[12] = {-0.375, -1.171875e-1, -1.025390625e-1, -1.4419555664063e-1,
                            -2.7757644653320e-1, -6.7659258842468e-1, -1.9935317337513,
                            -6.8839142681099, -2.7248827311269e01, -1.2159789187654e02,
                            -6.0384407670507e02, -3.3022722944809e03};
      const int kn = x >= 50.0 ? 7 : (x >= 35.0 ? 9 : 12);
      const double ca = exp(x) / sqrt(2.0 * kPi * x), xr = 1.0 / x;
      double p = 1.0;
      bi0 = 1.0;
      bi1 = 1.0;
      for (int k = 0; k < kn; ++k) {
        p *= xr;
        bi0 += a[k] * p;
        bi1 += b[k] * p;
      }
      bi0 *= ca;
      bi1 *= ca;
    }
    const double a1[8] = {0.125, 0.2109375, 1.0986328125, 1.1775970458984e01,
                          2.1461706161499e02, 5.9511522710323e03, 2.3347645606175e05,
                          1.2312234987631e07};
    const double xr2 = 1.0 / (x * x);
    double p = 1.0;
    k0v = 1.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      p *= xr2;
      k0v += a1[k] * p;
    }
    k0v = 0.5 / x * k0v / bi0;
  }
  bk0 = k0v;
  bk1 = (1.0 / x - bi1 * k0v) / bi0;
}

// the reference's K0/K1 (plain path) or the fast evaluation (mixed paths)
template <bool FAST>
HM_DEV void sf_k01_t(double x, double& bk0, double& bk1) {
  if (FAST) sf_k01_fast(x, bk0, bk1);
  else sf_k01(x, bk0, bk1);
}

// J0, J1, Y0, Y1 polynomial fits (JY01B :663; 1.0766115157 is a REAL(4) literal)
HM_DEV void sf_jy01(double x, double& j0, double& j1, double& y0, double& y1) {
  if (x == 0.0) { j0 = 1.0; j1 = 0.0; y0 = -1.0e300; y1 = -1.0e300; return; }
  if (x <= 4.0) {
    const double t = x / 4.0, t2 = t * t;
    j0 = ((((((-.5014415e-3 * t2 + .76771853e-2) * t2 - .0709253492) * t2 + .4443584263) * t2 -
            1.7777560599) * t2 + 3.9999973021) * t2 - 3.9999998721) * t2 + 1.0;
    j1 = t * (((((((-.1289769e-3 * t2 + .22069155e-2) * t2 - .0236616773) * t2 + .1777582922) *
                  t2 - .8888839649) * t2 + 2.6666660544) * t2 - 3.9999999710) * t2 + 1.9999999998);
    const double lg = 2.0 / kPi * log(x / 2.0);
    y0 = lg * j0 + ((((((((-.567433e-4 * t2 + .859977e-3) * t2 - .94855882e-2) * t2 + .0772975809) *
                         t2 - .4261737419) * t2 + 1.4216421221) * t2 - 2.3498519931) * t2 +
                      F32(1.0766115157)) * t2 + .3674669052);
    y1 = lg * j1 + ((((((((.6535773e-3 * t2 - .0108175626) * t2 + .107657606) * t2 - .7268945577) *
                         t2 + 3.1261399273) * t2 - 7.3980241381) * t2 + 6.8529236342) * t2 +
                      .3932562018) * t2 - .6366197726) / x;
    return;
  }
  const double t = 4.0 / x, t2 = t * t, a0 = sqrt(2.0 / (kPi * x));
  const double p0 = ((((-.9285e-5 * t2 + .43506e-4) * t2 - .122226e-3) * t2 + .434725e-3) * t2 -
                     .4394275e-2) * t2 + .999999997;
  const double q0 = t * (((((.8099e-5 * t2 - .35614e-4) * t2 + .85844e-4) * t2 - .218024e-3) * t2 +
                          .1144106e-2) * t2 - .031249995);
  const double p1 = ((((.10632e-4 * t2 - .50363e-4) * t2 + .145575e-3) * t2 - .559487e-3) * t2 +
                     .7323931e-2) * t2 + 1.000000004;
  const double q1 = t * (((((-.9173e-5 * t2 + .40658e-4) * t2 - .99941e-4) * t2 + .266891e-3) * t2 -
                          .1601836e-2) * t2 + .093749994);
  double s0, c0, s1, c1;
  sincos(x - .25 * kPi, &s0, &c0);
  sincos(x - .75 * kPi, &s1, &c1);
  j0 = a0 * (p0 * c0 - q0 * s0);
  y0 = a0 * (p0 * s0 + q0 * c0);
  j1 = a0 * (p1 * c1 - q1 * s1);
  y1 = a0 * (p1 * s1 + q1 * c1);
}

// Wynn epsilon limit of partial sums s[0..n] (Limes :370). s is overwritten;
// a is caller scratch of n+1 entries. The exact equality tests are kept: they
// decide the early exit exactly as in the reference.
HM_DEV double sf_limes(int n, double* s, double* a) {
  double rl = s[n], amax = 0.0;
  for (int i = 0; i <= n; ++i) a[i] = 0.0;
  for (int k = 1;; ++k) {
    if (n - 2 * k + 1 < 0) break;
    for (int i = 0; i <= n - 2 * k + 1; ++i)
      a[i] = (s[i + 1] != s[i]) ? a[i + 1] + 1.0 / (s[i + 1] - s[i]) : a[i + 1];
    if (n - 2 * k < 0) break;
    for (int i = 0; i <= n - 2 * k; ++i) {
      s[i] = (a[i + 1] != a[i]) ? s[i + 1] + 1.0 / (a[i + 1] - a[i]) : s[i + 1];
      if (fabs(a[i]) > amax) {
        amax = fabs(a[i]);
        rl = s[i];
        if (s[i + 1] == s[i]) return rl;
      }
    }
  }
  return rl;
}

// ---------------------------------------------------------------------------
// Finite depth: wave part FG = (G, dG/dR, dG/dzf) with 1/r + 1/r1 removed.
// ---------------------------------------------------------------------------
struct FinOut { double g_re, g_im, gr_re, gr_im, gz_re, gz_im; };

// propagating mode (shared head of EIGEN/EIGENE, FinGreen3D.f90:322-334)
HM_DEV void fin_propagating(double r, double zfh, double zph, double wk, double h, FinOut& o) {
  const double nm0 = h / 2.0 * (1.0 + sinh(2.0 * wk * h) / (2.0 * wk * h));
  const double p0 = kPi / nm0 * cosh(wk * zfh) * cosh(wk * zph);
  const double dp0 = kPi / nm0 * sinh(wk * zfh) * cosh(wk * zph);
  double j0, j1, y0, y1;
  sf_jy01(wk * r, j0, j1, y0, y1);
  o.g_re = p0 * -y0;          o.g_im = p0 * j0;
  o.gr_re = -wk * p0 * -y1;   o.gr_im = -wk * p0 * j1;
  o.gz_re = wk * dp0 * -y0;   o.gz_im = wk * dp0 * j0;
}

// remove 1/r + 1/r1 (common tail, FinGreen3D.f90:349-358)
HM_DEV void fin_strip(double r, double zf, double zp, FinOut& o) {
  const double sr = sqrt(r * r + (zf - zp) * (zf - zp));
  const double sr1 = sqrt(r * r + (zf + zp) * (zf + zp));
  o.g_re -= 1.0 / sr + 1.0 / sr1;
  o.gr_re += (r / sr) / (sr * sr) + (r / sr1) / (sr1 * sr1);
  o.gz_re += ((zf - zp) / sr) / (sr * sr) + ((zf + zp) / sr1) / (sr1 * sr1);
}

// one evanescent term of the eigenfunction series (EIGEN/EIGENE loop body);
// FAST: K0/K1 by sf_k01_fast (mixed-precision paths). The depth factors are
// multiplied together before the mode factor, and both come from sincos, so
// that the term of the mirrored pair (zfh, zph swapped) is the same number:
// a pair and its mirror evaluated one by one, or together by fin_evanescent_sym,
// give identical partial sums.
template <bool FAST = false>
HM_DEV void fin_evanescent(double km, double r, double zfh, double zph, double h, double& t1,
                           double& t2, double& t3) {
#pragma clang fp contract(off)
  const double twokmh = 2.0 * km * h;
  const double pm = 4.0 / (1.0 + sin(twokmh) / twokmh) / h;
  double k0, k1;
  sf_k01_t<FAST>(km * r, k0, k1);
  double skzf, ckzf, skzp, ckzp;
  sincos(km * zfh, &skzf, &ckzf);
  sincos(km * zph, &skzp, &ckzp);
  const double cc = ckzf * ckzp;
  t1 = (pm * k0) * cc;
  t2 = (-pm * km * k1) * cc;
  t3 = (-pm * km * k0) * (skzf * ckzp);
}

// fin_evanescent and t4, the derivative of the term with respect to zph: the
// mirrored pair's t3 (t1, t2 are symmetric in zfh, zph), bit for bit what
// fin_evanescent gives for the mirror. FAST as there.
template <bool FAST = false>
HM_DEV void fin_evanescent_sym(double km, double r, double zfh, double zph, double h, double& t1, double& t2,
                               double& t3, double& t4) {
#pragma clang fp contract(off)
  const double twokmh = 2.0 * km * h;
  const double pm = 4.0 / (1.0 + sin(twokmh) / twokmh) / h;
  double k0, k1;
  sf_k01_t<FAST>(km * r, k0, k1);
  double skzf, ckzf, skzp, ckzp;
  sincos(km * zfh, &skzf, &ckzf);
  sincos(km * zph, &skzp, &ckzp);
  const double cc = ckzf * ckzp;
  const double d0 = -pm * km * k0;
  t1 = (pm * k0) * cc;
  t2 = (-pm * km * k1) * cc;
  t3 = d0 * (skzf * ckzp);
  t4 = d0 * (skzp * ckzf);
}

// Evanescent terms summed directly, without the epsilon acceleration, for the
// cross approximation's entries between bodies (R/h >= 0.05; mixed path only).
// With rho = R/h and kappa_n = k_n h >= (n - 1/2) pi, the terms of the three
// series are bounded by (1/h^2) (5.72/rho) sqrt(pi x / 2) exp(-x) at
// x = kappa_n rho (pm_n <= 4.77/h for kappa_n >= pi/2 + pi, K0, K1 <= 1.2
// sqrt(pi/(2x)) exp(-x) for x >= 2), and the terms after n decrease at least
// by exp(-pi rho) each, so stopping where that bound over (1 - exp(-pi rho))
// falls below 1e-10 / h^2 leaves a tail below 1e-10 of the scale of dG/dR at
// R = h (and about 1e-12 of the entries themselves at R/h <= 0.5). The count
// comes from three fixed-point steps of x = ln(7.17 / (rho eps (1 - q))) +
// ln(x) / 2 and is capped by the 199 evanescent roots held per frequency.
HM_DEV int fin_direct_terms(double rr, double h) {
  const float rho = (float)(fabs(rr) / h);
  const float q = __expf(-3.14159265f * rho);
  const float c = __logf(7.17f / (rho * 1.0e-10f * (1.0f - q)));
  float x = 30.0f;
  for (int k = 0; k < 3; ++k) x = c + 0.5f * __logf(x);
  // kappa_N = x / rho >= (N - 1/2) pi: N = x / (pi rho) + 1/2, n = N + 1
  const int n = (int)(x / (3.14159265f * rho) + 1.5f) + 1;
  return n < 2 ? 2 : (n > 200 ? 200 : n);       // 200 roots per frequency: modes 1..199
}

// fin_evanescent with its depth factors and normalisation given: the same
// operations on the same values (pm as fin_evanescent forms it; skzf, ckzf the
// sine and cosine of km zfh, ckzp the cosine of km zph)
template <bool FAST = false>
HM_DEV void fin_evanescent_given(double km, double pm, double r, double skzf, double ckzf, double ckzp, double& t1,
                                 double& t2, double& t3) {
#pragma clang fp contract(off)
  double k0, k1;
  sf_k01_t<FAST>(km * r, k0, k1);
  const double cc = ckzf * ckzp;
  t1 = (pm * k0) * cc;
  t2 = (-pm * km * k1) * cc;
  t3 = (-pm * km * k0) * (skzf * ckzp);
}

// the evanescent sums of modes 1 .. n-1 (n = fin_direct_terms), FAST K0/K1,
// with the depth factors of source panel j and field point i and the modes'
// normalisations from per-frequency tables (evt: sin, cos of k_m (z + h) per
// mode and point, leading dimension ld; pmt: pm per mode)
HM_DEV void fin_evanescent_direct_tab(double r, const double* wvn, const double* evt, long long ld,
                                      const double* pmt, int i, int j, int n, double& s1, double& s2, double& s3) {
  s1 = s2 = s3 = 0.0;
  for (int m = 1; m < n; ++m) {
    const double* ej = evt + 2 * ((long long)m * ld + j);
    const double ckzp = evt[2 * ((long long)m * ld + i) + 1];
    double t1, t2, t3;
    fin_evanescent_given<true>(wvn[m], pmt[m], r, ej[0], ej[1], ckzp, t1, t2, t3);
    s1 += t1;
    s2 += t2;
    s3 += t3;
  }
}

// the part of fin_evanescent_direct_tab over modes m0, m0 + step, ... < n
HM_DEV void fin_evanescent_direct_tab_part(double r, const double* wvn, const double* evt, long long ld,
                                           const double* pmt, int i, int j, int n, int m0, int step, double& s1,
                                           double& s2, double& s3) {
  s1 = s2 = s3 = 0.0;
  for (int m = m0; m < n; m += step) {
    const double* ej = evt + 2 * ((long long)m * ld + j);
    const double ckzp = evt[2 * ((long long)m * ld + i) + 1];
    double t1, t2, t3;
    fin_evanescent_given<true>(wvn[m], pmt[m], r, ej[0], ej[1], ckzp, t1, t2, t3);
    s1 += t1;
    s2 += t2;
    s3 += t3;
  }
}

// the evanescent sums of modes 1 .. n-1 (n = fin_direct_terms), FAST K0/K1
HM_DEV void fin_evanescent_direct(double r, double zfh, double zph, const double* wvn, double h, int n,
                                  double& s1, double& s2, double& s3) {
  s1 = s2 = s3 = 0.0;
  for (int i = 1; i < n; ++i) {
    double t1, t2, t3;
    fin_evanescent<true>(wvn[i], r, zfh, zph, h, t1, t2, t3);
    s1 += t1;
    s2 += t2;
    s3 += t3;
  }
}

// R/h >= 0.5: at most 9 evanescent terms, early exit at 1e-6 (EIGEN :281)
HM_DEV void fin_eigen(double r, double zf, double zp, const double* wvn, double h, FinOut& o) {
  const double zfh = zf + h, zph = zp + h;
  fin_propagating(r, zfh, zph, wvn[0], h, o);
  for (int i = 1; i < 10; ++i) {
    double t1, t2, t3;
    fin_evanescent(wvn[i], r, zfh, zph, h, t1, t2, t3);
    o.g_re += t1; o.gr_re += t2; o.gz_re += t3;
    if (fabs(t1) < F32(1.e-6) && fabs(t2) < F32(1.e-6) && fabs(t3) < F32(1.e-6)) break;
  }
  fin_strip(r, zf, zp, o);
}

// The evanescent-mode sums alone (what EIGEN / EIGENE add to the propagating
// mode before the 1/r + 1/r1 strip): the part a learned surrogate replaces.
HM_DEV void fin_eigen_evan(double r, double zf, double zp, const double* wvn, double h, double& tg, double& tr,
                           double& tz) {
  const double zfh = zf + h, zph = zp + h;
  tg = tr = tz = 0.0;
  for (int i = 1; i < 10; ++i) {
    double t1, t2, t3;
    fin_evanescent(wvn[i], r, zfh, zph, h, t1, t2, t3);
    tg += t1; tr += t2; tz += t3;
    if (fabs(t1) < F32(1.e-6) && fabs(t2) < F32(1.e-6) && fabs(t3) < F32(1.e-6)) break;
  }
}

// Series lengths (FinGreen3D.f90:433, :597). Like every discrete decision of the
// reference (regime, series length, near-panel flag), they are evaluated without
// multiply-add contraction so that they round exactly as the reference does.
HM_DEV int fin_eigene_terms(double rr, double h) {
#pragma clang fp contract(off)
  return (int)(F32(-88.89) * fabs(rr / h) + F32(54.45));
}
HM_DEV int fin_pidcock_terms(double rr, double h) {
#pragma clang fp contract(off)
  return (int)(F32(-1010.10) * fabs(rr / h) + F32(100.50));
}
// horizontal distance R of source (xf,yf) and field point (xp,yp)
HM_DEV double horiz_dist(double xf, double xp, double yf, double yp) {
#pragma clang fp contract(off)
  const double xpq = xf - xp, ypq = yf - yp;
  return sqrt(xpq * xpq + ypq * ypq);
}
// near-panel flag DIST <= 50*PNSZ (CalGreenFuncMulti.f90:134-139)
HM_DEV int near_flag(const double* xp, const double* xq, double pnsz) {
#pragma clang fp contract(off)
  const double d0 = xp[0] - xq[0], d1 = xp[1] - xq[1], d2 = xp[2] - xq[2];
  return sqrt(d0 * d0 + d1 * d1 + d2 * d2) <= 50.0 * pnsz ? 1 : 0;
}

// 0.05 <= R/h < 0.5: NT-1 evanescent terms then Wynn epsilon (EIGENE :413)
HM_DEV void fin_eigene(double r, double zf, double zp, const double* wvn, double h, FinOut& o) {
  double s1[kEigeneMaxTerms], s2[kEigeneMaxTerms], s3[kEigeneMaxTerms], sa[kEigeneMaxTerms];
  const int nt = fin_eigene_terms(r, h);
  const double zfh = zf + h, zph = zp + h;
  fin_propagating(r, zfh, zph, wvn[0], h, o);
  s1[0] = s2[0] = s3[0] = 0.0;
  for (int i = 1; i < nt; ++i) {
    double t1, t2, t3;
    fin_evanescent(wvn[i], r, zfh, zph, h, t1, t2, t3);
    s1[i] = s1[i - 1] + t1;
    s2[i] = s2[i - 1] + t2;
    s3[i] = s3[i - 1] + t3;
  }
  o.g_re += sf_limes(nt - 1, s1, sa);
  o.gr_re += sf_limes(nt - 1, s2, sa);
  o.gz_re += sf_limes(nt - 1, s3, sa);
  fin_strip(r, zf, zp, o);
}

HM_DEV void fin_eigene_evan(double r, double zf, double zp, const double* wvn, double h, double& tg, double& tr,
                            double& tz) {
  double s1[kEigeneMaxTerms], s2[kEigeneMaxTerms], s3[kEigeneMaxTerms], sa[kEigeneMaxTerms];
  const int nt = fin_eigene_terms(r, h);
  const double zfh = zf + h, zph = zp + h;
  s1[0] = s2[0] = s3[0] = 0.0;
  for (int i = 1; i < nt; ++i) {
    double t1, t2, t3;
    fin_evanescent(wvn[i], r, zfh, zph, h, t1, t2, t3);
    s1[i] = s1[i - 1] + t1;
    s2[i] = s2[i - 1] + t2;
    s3[i] = s3[i - 1] + t3;
  }
  tg = sf_limes(nt - 1, s1, sa);
  tr = sf_limes(nt - 1, s2, sa);
  tz = sf_limes(nt - 1, s3, sa);
}

// Rankine image sums by the Chebyshev fit (IMGS/DGSR/DGSZ, FinGreen3D.f90:941-1114)
HM_DEV void fin_images(double r, double zf, double zp, double h, double& g, double& gr, double& gz) {
  const double amn[25] = {0.0, -0.02525711, 0.00086546, -0.00004063, 0.00000193,
                          0.05051418, -0.00692380, 0.00073292, -0.00006636, 0.00000398,
                          0.00230838, -0.00097875, 0.00020597, -0.00003333, 0.00000524,
                          0.00012934, -0.00010879, 0.00003965, -0.00000891, 0.0,
                          0.00000913, -0.00001270, 0.00000466, 0.0, 0.0};
  const double t1 = r / h, t2 = (zf - zp) / h, t3 = (zf + zp + 2.0 * h) / h;
  double s1 = 0.0, s1r = 0.0, s1z = 0.0;
  for (int l = -1; l <= 1; ++l) {
    const double u = t2 + 2.0 * l, w = t3 + 2.0 * l;
    const double a = sqrt(t1 * t1 + u * u), b = sqrt(t1 * t1 + w * w);
    s1 += 1.0 / a + 1.0 / b;
    s1r += t1 / (a * a * a) + t1 / (b * b * b);
    s1z += u / (a * a * a) + w / (b * b * b);
  }
  // powers t^(2m) and t^(2n), t^(2n-1) by products instead of pow()
  double p1[5], q2[5], q3[5];
  p1[0] = q2[0] = q3[0] = 1.0;
  for (int k = 1; k < 5; ++k) {
    p1[k] = p1[k - 1] * t1 * t1;
    q2[k] = q2[k - 1] * t2 * t2;
    q3[k] = q3[k - 1] * t3 * t3;
  }
  double s2 = 0.0, s2r = 0.0, s2z = 0.0;
  for (int m = 0; m < 5; ++m)
    for (int n = 0; n < 5; ++n) {
      const double c = amn[m + 5 * n];
      s2 += c * p1[m] * (q2[n] + q3[n]);
      if (m >= 1) s2r += 2.0 * m * c * (p1[m - 1] * t1) * (q2[n] + q3[n]);
      if (n >= 1) s2z += 2.0 * n * c * p1[m] * (q2[n - 1] * t2 + q3[n - 1] * t3);
    }
  g = (s1 - 2.0 + s2) / h;
  gr = (-s1r + s2r) / (h * h);
  gz = (-s1z + s2z) / (h * h);
}

// 0.0005 <= R/h < 0.05: accelerated series + Wynn epsilon + image sums (PIDCOCK :558).
// Split into the part before the series, one series term, and the part after
// the epsilon limits, so that the lane-cooperative assembly kernel can share them.
HM_DEV void fin_pidcock_head(double r, double zf, double zp, const double* wvn, double h, FinOut& o) {
  const double wk = wvn[0], zfh = zf + h, zph = zp + h;
  const double nm0 = h / 2.0 * (1.0 + sinh(2.0 * wk * h) / (2.0 * wk * h));
  const double p0 = kPi / nm0 * cosh(wk * zfh) * cosh(wk * zph);
  const double dp0 = wk * kPi / nm0 * sinh(wk * zfh) * cosh(wk * zph);
  double j0, j1, y0, y1;
  sf_jy01(wk * r, j0, j1, y0, y1);
  o.g_re = p0 * -y0 + 2.0 / h * (0.5772156649015328 + log(r / 4.0 / h));
  o.g_im = p0 * j0;
  o.gr_re = -wk * p0 * -y1 + 2.0 / (r * h);
  o.gr_im = -wk * p0 * j1;
  o.gz_re = dp0 * -y0;
  o.gz_im = dp0 * j0;
}
// term i >= 1 of the three accelerated series (mode i minus its large-i form);
// FAST: K0/K1 by sf_k01_fast (mixed-precision paths). As in fin_evanescent the
// depth factors are multiplied together first, every sine and cosine comes
// from sincos and nothing is contracted, so that the mirrored pair's terms
// (fin_pidcock_term_sym's t4, and the cached form) are the same numbers.
template <bool FAST = false>
HM_DEV void fin_pidcock_term(int i, double um, double r, double zfh, double zph, double v, double h, double& t1,
                             double& t2, double& t3) {
#pragma clang fp contract(off)
  const double umt = i * kPi / h;
  double k0, k1, k0t, k1t;
  sf_k01_t<FAST>(um * r, k0, k1);
  sf_k01_t<FAST>(umt * r, k0t, k1t);
  const double ac = (um * um + v * v) / ((um * um + v * v) * h - v);
  double su, cu, sut, cut, sup, cup, sutp, cutp;
  sincos(um * zfh, &su, &cu);
  sincos(umt * zfh, &sut, &cut);
  sincos(um * zph, &sup, &cup);
  sincos(umt * zph, &sutp, &cutp);
  const double am = ac * (cu * cup);
  const double amt = (cut * cutp) / h;
  const double dam = (-um * ac) * (su * cup);
  const double damt = (-umt * (sut * cutp)) / h;
  t1 = am * k0 - amt * k0t;
  t2 = -um * am * k1 + umt * amt * k1t;
  t3 = dam * k0 - damt * k0t;
}
// fin_pidcock_term and t4, the derivative of the term with respect to zph:
// the term of the mirrored pair (field and source depths swapped) has t1, t2
// unchanged and t4 as its t3 (the series is symmetric in zfh, zph), bit for
// bit. FAST as in fin_pidcock_term.
template <bool FAST = false>
HM_DEV void fin_pidcock_term_sym(int i, double um, double r, double zfh, double zph, double v, double h, double& t1,
                                 double& t2, double& t3, double& t4) {
#pragma clang fp contract(off)
  const double umt = i * kPi / h;
  double k0, k1, k0t, k1t;
  sf_k01_t<FAST>(um * r, k0, k1);
  sf_k01_t<FAST>(umt * r, k0t, k1t);
  const double ac = (um * um + v * v) / ((um * um + v * v) * h - v);
  double su, cu, sut, cut, sup, cup, sutp, cutp;
  sincos(um * zfh, &su, &cu);
  sincos(umt * zfh, &sut, &cut);
  sincos(um * zph, &sup, &cup);
  sincos(umt * zph, &sutp, &cutp);
  const double am = ac * (cu * cup);
  const double amt = (cut * cutp) / h;
  const double dam = (-um * ac) * (su * cup);
  const double damt = (-umt * (sut * cutp)) / h;
  const double damp = (-um * ac) * (sup * cu);
  const double damtp = (-umt * (sutp * cut)) / h;
  t1 = am * k0 - amt * k0t;
  t2 = -um * am * k1 + umt * amt * k1t;
  t3 = dam * k0 - damt * k0t;
  t4 = damp * k0 - damtp * k0t;
}
// The frequency-independent half of the PIDCOCK term of row i (the i pi / h
// mode): c[0] = amt k0t, c[1] = umt amt k1t, c[2] = damt k0t, c[3] = damtp k0t
// (depends on R, the depths and h only; cached once per geometry), with the
// operations of fin_pidcock_term_sym
HM_DEV void fin_pidcock_tilde(int i, double r, double zfh, double zph, double h, double* c) {
#pragma clang fp contract(off)
  const double umt = i * kPi / h;
  double k0t, k1t;
  sf_k01(umt * r, k0t, k1t);
  double sut, cut, sutp, cutp;
  sincos(umt * zfh, &sut, &cut);
  sincos(umt * zph, &sutp, &cutp);
  const double amt = (cut * cutp) / h;
  const double damt = (-umt * (sut * cutp)) / h;
  const double damtp = (-umt * (sutp * cut)) / h;
  c[0] = amt * k0t;
  c[1] = umt * amt * k1t;
  c[2] = damt * k0t;
  c[3] = damtp * k0t;
}
// fin_pidcock_term_sym with the i pi / h half from the cache (the same numbers)
HM_DEV void fin_pidcock_term_cached(double um, double r, double zfh, double zph, double v, double h, const double* c,
                                    double& t1, double& t2, double& t3, double& t4) {
#pragma clang fp contract(off)
  double k0, k1;
  sf_k01(um * r, k0, k1);
  const double ac = (um * um + v * v) / ((um * um + v * v) * h - v);
  double su, cu, sup, cup;
  sincos(um * zfh, &su, &cu);
  sincos(um * zph, &sup, &cup);
  const double am = ac * (cu * cup);
  const double dam = (-um * ac) * (su * cup);
  const double damp = (-um * ac) * (sup * cu);
  t1 = am * k0 - c[0];
  t2 = -um * am * k1 + c[1];
  t3 = dam * k0 - c[2];
  t4 = damp * k0 - c[3];
}
HM_DEV void fin_pidcock_tail(double l1, double l2, double l3, double r, double zf, double zp, double h, FinOut& o) {
  double ig, igr, igz;
  fin_images(r, zf, zp, h, ig, igr, igz);
  o.g_re += 4.0 * l1 + ig;
  o.gr_re += 4.0 * l2 + igr;
  o.gz_re += 4.0 * l3 + igz;
  fin_strip(r, zf, zp, o);
}
HM_DEV void fin_pidcock(double r, double zf, double zp, double v, const double* wvn, double h,
                        FinOut& o) {
  double s1[kPidcockMaxTerms], s2[kPidcockMaxTerms], s3[kPidcockMaxTerms], sa[kPidcockMaxTerms];
  const double zfh = zf + h, zph = zp + h;
  fin_pidcock_head(r, zf, zp, wvn, h, o);
  const int nt = fin_pidcock_terms(r, h);
  s1[0] = s2[0] = s3[0] = 0.0;
  for (int i = 1; i < nt; ++i) {
    double t1, t2, t3;
    fin_pidcock_term(i, wvn[i], r, zfh, zph, v, h, t1, t2, t3);
    s1[i] = s1[i - 1] + t1;
    s2[i] = s2[i - 1] + t2;
    s3[i] = s3[i - 1] + t3;
  }
  const double l1 = sf_limes(nt - 1, s1, sa);
  const double l2 = sf_limes(nt - 1, s2, sa);
  const double l3 = sf_limes(nt - 1, s3, sa);
  fin_pidcock_tail(l1, l2, l3, r, zf, zp, h, o);
}

// ---- Linton regime (R/h < 0.0005, FinGreen3D.f90:712-861) -------------------

// normalisation of mode km (NM :2289); for the propagating mode km = -i*wk it is real
HM_DEV double lin_nm_real(double h, double km) { return h / 2.0 * (1.0 + sin(2.0 * km * h) / (2.0 * km * h)); }
HM_DEV double lin_nm_prop(double h, double wk) { return h / 2.0 * (1.0 + sinh(2.0 * wk * h) / (2.0 * wk * h)); }

// quadrature integrands: kind 0 = G1/G2/G3 (tag 1..3), kind 1 = F1/F2 (tag 1..2)
struct QuadFn {
  int kind, tag, no;
  double v, r, zf, zp, h, wvn;
  HM_DEV double operator()(double t) const {
    if (fabs(t) < 1.0e-8) return 0.0;
    if (kind == 1) {
      const double e = (no == 1) ? exp(-r * r / 4.0 / t + wvn * wvn * t) : exp(-r * r / 4.0 / t - wvn * wvn * t);
      return tag == 1 ? e / 4.0 / kPi / t : -r * e / 8.0 / kPi / (t * t);
    }
    const double val[4] = {-zp - zf, 2.0 * h - zp + zf, 2.0 * h + zp - zf, 4.0 * h + zp + zf};
    const double dv[4] = {-1.0, 1.0, -1.0, 1.0};
    double sum = 0.0;
    for (int i = 0; i < 4; ++i) {
      const double ev = exp(-v * val[i]), arg = val[i] / 2.0 / t - v * t;
      if (tag == 3) sum = sum - v * dv[i] * ev * sf_erfc(arg) + ev * sf_derfc(arg) * dv[i] / 2.0 / t;
      else sum = sum + ev * sf_erfc(arg);
    }
    const double e = exp(v * v * t * t - r * r / 4.0 / (t * t));
    return tag == 2 ? r / 2.0 / (t * t * t) * e * sum : e * sum / t;
  }
};

// adaptive G7-K15 bisection (GK_INTEG*/AQUAD*, FinGreen3D.f90:2387-2745) with an
// explicit stack; leaves are summed left to right
HM_DEV double lin_quad(const QuadFn& f, double a0, double b0, double eps) {
  const double xg[7] = {-0.949107912342759, -0.741531185599394, -0.405845151377397, 0.0,
                        0.405845151377397, 0.741531185599394, 0.949107912342759};
  const double wg[7] = {0.129484966168870, 0.279705391489277, 0.381830050505119, 0.417959183673469,
                        0.381830050505119, 0.279705391489277, 0.129484966168870};
  const double xk[8] = {-0.991455371120813, -0.864864423359769, -0.586087235467691,
                        -0.207784955007898, 0.207784955007898, 0.586087235467691,
                        0.864864423359769, 0.991455371120813};
  const double wk[15] = {0.022935322010529, 0.063092092629979, 0.104790010322250,
                         0.140653259715525, 0.169004726639267, 0.190350578064785,
                         0.204432940075298, 0.209482141084728, 0.204432940075298,
                         0.190350578064785, 0.169004726639267, 0.140653259715525,
                         0.104790010322250, 0.063092092629979, 0.022935322010529};
  double sa[64], sb[64];
  int top = 0;
  sa[0] = a0; sb[0] = b0;
  double total = 0.0;
  while (top >= 0) {
    const double a = sa[top], b = sb[top];
    --top;
    double fg[7], g = 0.0, k = 0.0;
    for (int i = 0; i < 7; ++i) {
      fg[i] = f(1 / 2.0 * (b + a + (b - a) * xg[i]));
      g = g + wg[i] * (b - a) / 2.0 * fg[i];
    }
    for (int i = 0; i < 15; ++i) {
      const double fx = (i & 1) ? fg[i >> 1] : f(1 / 2.0 * (b + a + (b - a) * xk[i >> 1]));
      k = k + wk[i] * (b - a) / 2.0 * fx;
    }
    const double err = pow(200.0 * fabs(k - g), 1.5);
    if (err > eps && fabs(a - b) > eps && top < 62) {
      ++top; sa[top] = (a + b) / 2.0; sb[top] = b;     // right half, processed second
      ++top; sa[top] = a; sb[top] = (a + b) / 2.0;     // left half, processed first
    } else {
      total += k;
    }
  }
  return total;
}

// small-R series of the expansion coefficients (SA0/DSA0/SAM/DSAM :1389-1694)
HM_DEV double lin_sa0(double r, double pa, double wk) {
  const double ga = 0.5772156649015328;
  double rm = 100.0, s = -ga - 2.0 * log(wk * pa / 2.0), sgn = -1.0, facn = 1.0, psi = 0.0;
  for (int n = 1; fabs(rm) > F32(1.e-6); ++n) {
    const double en = sf_en(n + 1, (r / pa) * (r / pa));
    psi += 1.0 / n;
    rm = sgn * pow(r / pa, 2.0 * n) / facn / n - pow(wk * pa / 2.0, 2.0 * n) / facn * en -
         2.0 * sgn / (facn * facn) * pow(wk * r / 2.0, 2.0 * n) * (log(wk * r / 2.0) + ga - psi);
    s += rm;
    sgn = -sgn;
    facn *= (n + 1);
  }
  return s / 4.0 / kPi;
}
HM_DEV double lin_dsa0(double r, double pa, double wk) {
  const double ga = 0.5772156649015328;
  double rm = 100.0, s = 0.0, sgn = -1.0, facn = 1.0, psi = 0.0;
  for (int n = 1; fabs(rm) > F32(1.e-6); ++n) {
    const double en = sf_en(n, (r / pa) * (r / pa));
    psi += 1.0 / n;
    rm = 2.0 * sgn * pow(r / pa, 2.0 * n - 1.0) / facn / pa +
         pow(wk * pa / 2.0, 2.0 * n) / facn * en * 2.0 * r / (pa * pa) -
         2.0 * sgn * n * wk / (facn * facn) * pow(wk * r / 2.0, 2.0 * n - 1.0) * (log(wk * r / 2.0) + ga - psi) -
         2.0 * sgn / (facn * facn) * pow(wk * r / 2.0, 2.0 * n) / r;
    s += rm;
    sgn = -sgn;
    facn *= (n + 1);
  }
  return s / 4.0 / kPi;
}
HM_DEV double lin_sam(double r, double pa, double km) {
  double rm = 100.0, s = 0.0, sgn = 1.0, facn = 1.0;
  const double x = (km * pa / 2.0) * (km * pa / 2.0);
  for (int n = 0; fabs(rm) > F32(1.e-6); ++n) {
    rm = sgn * pow(r / pa, 2.0 * n) / facn * sf_en(n + 1, x);
    s += rm;
    sgn = -sgn;
    facn *= (n + 1);
  }
  return s / 4.0 / kPi;
}
HM_DEV double lin_dsam(double r, double pa, double km) {
  double rm = 100.0, s = 0.0, sgn = -1.0, facn = 1.0;
  const double x = (km * pa / 2.0) * (km * pa / 2.0);
  for (int n = 1; fabs(rm) > F32(1.e-6); ++n) {
    rm = 2.0 * n * sgn * pow(r / pa, 2.0 * n - 1.0) / facn / pa * sf_en(n + 1, x);
    s += rm;
    sgn = -sgn;
    facn *= (n + 1);
  }
  return s / 4.0 / kPi;
}

// expansion coefficient i (0 = propagating) and its R-derivative (COEF/DCOEF :1168-1340)
HM_DEV void lin_coef(int i, double r, double pa, double a, const double* wvn, double& c, double& dc) {
  const double eps = F32(1.e-6), dnt = pa * pa / 4.0, wk = wvn[0];
  if (fabs(r) < F32(1.e-6)) {
    if (i == 0) { c = -sf_ei(wk * wk * dnt) / 4.0 / kPi; dc = 0.0; }
    else { c = sf_e1(wvn[i] * wvn[i] * dnt) / 4.0 / kPi; dc = 0.0; }
    return;
  }
  if (i == 0) {
    if (r / a <= 0.5) { c = lin_sa0(r, pa, wk); dc = lin_dsa0(r, pa, wk); return; }
    double j0, j1, y0, y1;
    sf_jy01(wk * r, j0, j1, y0, y1);
    QuadFn f1{1, 1, 1, 0.0, r, 0.0, 0.0, 0.0, wk}, f2{1, 2, 1, 0.0, r, 0.0, 0.0, 0.0, wk};
    c = -y0 / 4.0 - lin_quad(f1, 0.0, dnt, eps);
    dc = wk * y1 / 4.0 - lin_quad(f2, 0.0, dnt, eps);
    return;
  }
  const double km = wvn[i];
  if (r / a <= 1.0) { c = lin_sam(r, pa, km); dc = lin_dsam(r, pa, km); return; }
  QuadFn f1{1, 1, i + 1, 0.0, r, 0.0, 0.0, 0.0, km}, f2{1, 2, i + 1, 0.0, r, 0.0, 0.0, 0.0, km};
  double p1 = dnt + 100.0;
  while (fabs(f1(p1)) > eps) p1 += 100.0;
  c = lin_quad(f1, dnt, p1, eps);
  p1 = dnt + 100.0;
  while (fabs(f2(p1)) > eps) p1 += 100.0;
  dc = lin_quad(f2, dnt, p1, eps);
}

HM_DEV void fin_linton(double r, double zf, double zp, double v, const double* wvn, double h, FinOut& o) {
  const double eps = F32(1.e-6), wk = wvn[0];
  double a;
  if (v < 0.02) a = 3324.2 * (v * v * v) - 258.89 * (v * v) + 8.3958 * v + 1.3208;
  else if (v < 0.1) a = -30.303 * (v * v * v) - 2.8138 * (v * v) + 2.7740 * v + 0.05052;
  else if (v < 1.0) a = 1.5699 * (v * v * v) - 2.5075 * (v * v) + 1.8316 * v + 0.1077;
  else a = -0.000069641 * (v * v * v) - 0.0017082 * (v * v) + 0.45566 * v + 0.80551;
  a = fabs(a) / (wk * h);
  const double pa = a * h;
  const double zfh = zf + h, zph = zp + h;
  // modal sum (the propagating term is real: cos(-i k z) = cosh(k z))
  double g = 0.0, gr = 0.0, gz = 0.0;
  for (int i = 0; i < 20; ++i) {
    double c, dc;
    lin_coef(i, r, pa, a, wvn, c, dc);
    double t1, t2, t3;
    if (i == 0) {
      const double nm = lin_nm_prop(h, wk), c1 = cosh(wk * zfh), c2 = cosh(wk * zph);
      t1 = -c / nm * c1 * c2;
      t2 = -dc / nm * c1 * c2;
      t3 = -(c * wk) / nm * sinh(wk * zfh) * c2;
    } else {
      const double km = wvn[i], nm = lin_nm_real(h, km), c2 = cos(km * zph);
      double s1, c1;
      sincos(km * zfh, &s1, &c1);
      t1 = -c / nm * c1 * c2;
      t2 = -dc / nm * c1 * c2;
      t3 = c * km / nm * s1 * c2;
    }
    g += t1; gr += t2; gz += t3;
    if (i > 1 && fabs(t1) < eps && fabs(t2) < eps && fabs(t3) < eps) break;
  }
  // error-function image terms
  const double val[4] = {-zp - zf, 2.0 * h - zp + zf, 2.0 * h + zp - zf, 4.0 * h + zp + zf};
  double e0 = 0.0, e1 = 0.0, e2 = 0.0;
  for (int i = 0; i < 4; ++i) {
    const double ro = sqrt(r * r + val[i] * val[i]);
    const double dror = r / ro, droz = val[i] * ((i & 1) ? 1.0 : -1.0) / ro;
    const double ef = sf_erfc(ro / pa), def = sf_derfc(ro / pa);
    e0 += ef / (4.0 * kPi * ro);
    e1 += (def / pa * dror * ro - dror * ef) / (4.0 * kPi * ro * ro);
    e2 += (def / pa * droz * ro - droz * ef) / (4.0 * kPi * ro * ro);
  }
  const double sr = sqrt(r * r + (zf - zp) * (zf - zp));
  const double sr1 = sqrt(r * r + (2.0 * h + zf + zp) * (2.0 * h + zf + zp));
  const double sr2 = sqrt(r * r + (zf + zp) * (zf + zp));
  const double dsr1r = r / sr1, dsr1z = (2.0 * h + zf + zp) / sr1, dsr2r = r / sr2, dsr2z = (zf + zp) / sr2;
  const double q1 = sf_erfc(sr1 / pa), dq1 = sf_derfc(sr1 / pa);
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;          // the 1/r term, present when r > 1e-6
  if (fabs(sr) > F32(1.e-6)) {
    const double dsrr = r / sr, dsrz = (zf - zp) / sr, q = sf_erfc(sr / pa), dq = sf_derfc(sr / pa);
    a0 = (q - 1.0) / (4.0 * kPi * sr);
    a1 = -(q - 1.0) / (4.0 * kPi * (sr * sr)) * dsrr + dq / (4.0 * kPi * sr * pa) * dsrr;
    a2 = -(q - 1.0) / (4.0 * kPi * (sr * sr)) * dsrz + dq / (4.0 * kPi * sr * pa) * dsrz;
  }
  e0 = e0 + a0 + q1 / (4.0 * kPi * sr1) - 1.0 / (4.0 * kPi * sr2);
  e1 = e1 + a1 - q1 / (4.0 * kPi * (sr1 * sr1)) * dsr1r + dq1 / (4.0 * kPi * sr1 * pa) * dsr1r + dsr2r / (4.0 * kPi * (sr2 * sr2));
  e2 = e2 + a2 - q1 / (4.0 * kPi * (sr1 * sr1)) * dsr1z + dq1 / (4.0 * kPi * sr1 * pa) * dsr1z + dsr2z / (4.0 * kPi * (sr2 * sr2));
  // free-surface integrals
  QuadFn q1f{0, 1, 0, v, r, zf, zp, h, 0.0}, q2f{0, 2, 0, v, r, zf, zp, h, 0.0}, q3f{0, 3, 0, v, r, zf, zp, h, 0.0};
  g = g - e0 - v / 2.0 / kPi * lin_quad(q1f, 0.0, pa / 2.0, eps);
  gr = gr - e1 + v / 2.0 / kPi * lin_quad(q2f, 0.0, pa / 2.0, eps);
  gz = gz - e2 - v / 2.0 / kPi * lin_quad(q3f, 0.0, pa / 2.0, eps);
  // propagating Bessel-J part (imaginary)
  double bj0, bj1, y0, y1;
  sf_jy01(wk * r, bj0, bj1, y0, y1);
  const double nm = lin_nm_prop(h, wk), chf = cosh(wk * zfh), chp = cosh(wk * zph);
  const double gi = -1.0 / 4.0 / nm * bj0 * chf * chp;
  const double gri = wk / 4.0 / nm * bj1 * chf * chp;
  const double gzi = -wk / 4.0 / nm * bj0 * sinh(wk * zfh) * chp;
  const double s = -4.0 * kPi;
  o.g_re = s * g;   o.g_im = s * gi;
  o.gr_re = s * gr; o.gr_im = s * gri;
  o.gz_re = s * gz; o.gz_im = s * gzi;
}

// Regime code of a pair from R/h (FINGREEN3D_R, FinGreen3D.f90:183-191)
enum Regime { kLinton = 0, kPidcock = 1, kEigene = 2, kEigen = 3 };
HM_DEV int fin_regime(double rr, double h) {
  const double rhp = fabs(rr / h);
  return rhp < 0.0005 ? kLinton : (rhp < 0.05 ? kPidcock : (rhp < 0.5 ? kEigene : kEigen));
}

// Assemble GRN(1..4) = (G, dG/dxf, dG/dyf, dG/dzf) from the R-form and add the
// flag-dependent Rankine terms (FINGREEN3D :81-155). V=0 / V=-1 limits included.
struct Grn4 { double re[4], im[4]; };

HM_DEV void fin_to_grn(const FinOut& fg, double xpq, double ypq, double zf, double zp,
                       double rr, bool wave, int flag, Grn4& g) {
  const double zpq = zf + zp;
  const double r = sqrt(rr * rr + (zp - zf) * (zp - zf)), r1 = sqrt(rr * rr + zpq * zpq);
  const double r13 = r1 * r1 * r1, r3 = r * r * r;
  for (int k = 0; k < 4; ++k) g.re[k] = g.im[k] = 0.0;
  if (wave) {
    if (rr > F32(1.e-6)) {
      g.re[1] = fg.gr_re * xpq / rr; g.im[1] = fg.gr_im * xpq / rr;
      g.re[2] = fg.gr_re * ypq / rr; g.im[2] = fg.gr_im * ypq / rr;
    }
    g.re[0] = fg.g_re; g.im[0] = fg.g_im;
    g.re[3] = fg.gz_re; g.im[3] = fg.gz_im;
  }
  if (flag == 1) {
    g.re[0] += 1.0 / r1;
    g.re[1] -= xpq / r13;
    g.re[2] -= ypq / r13;
    g.re[3] -= zpq / r13;
  } else {
    g.re[0] += 1.0 / r1 + 1.0 / r;
    g.re[1] -= xpq / r13 + xpq / r3;
    g.re[2] -= ypq / r13 + ypq / r3;
    g.re[3] -= zpq / r13 + (zf - zp) / r3;
  }
}

HM_DEV void fin_limit_inf(double xpq, double ypq, double zf, double zp, double rr, int flag, Grn4& g) {
  // V = -1 (infinite frequency): the image enters with the opposite sign
  const double zpq = zf + zp;
  const double r = sqrt(rr * rr + (zp - zf) * (zp - zf)), r1 = sqrt(rr * rr + zpq * zpq);
  const double r13 = r1 * r1 * r1, r3 = r * r * r;
  for (int k = 0; k < 4; ++k) g.re[k] = g.im[k] = 0.0;
  if (flag == 1) {
    g.re[0] = -1.0 / r1; g.re[1] = xpq / r13; g.re[2] = ypq / r13; g.re[3] = zpq / r13;
  } else {
    g.re[0] = -1.0 / r1 + 1.0 / r;
    g.re[1] = xpq / r13 - xpq / r3;
    g.re[2] = ypq / r13 - ypq / r3;
    g.re[3] = zpq / r13 - (zf - zp) / r3;
  }
}

// Full FINGREEN3D for one pair, dispatching on the regime (used by the parity
// entry point; the assembly launches per-regime kernels instead).
HM_DEV void fingreen3d(double xf, double xp, double yf, double yp, double zf, double zp, double v,
                       const double* wvn, double h, int flag, Grn4& g) {
  const double xpq = xf - xp, ypq = yf - yp;
  const double rr = horiz_dist(xf, xp, yf, yp);
  const bool zero = fabs(v) < F32(1.e-8);
  if (!zero && fabs(v + 1.0) < F32(1.e-8)) { fin_limit_inf(xpq, ypq, zf, zp, rr, flag, g); return; }
  FinOut fg{0, 0, 0, 0, 0, 0};
  if (!zero) {
    switch (fin_regime(rr, h)) {
      case kLinton: fin_linton(rr, zf, zp, v, wvn, h, fg); break;
      case kPidcock: fin_pidcock(rr, zf, zp, v, wvn, h, fg); break;
      case kEigene: fin_eigene(rr, zf, zp, wvn, h, fg); break;
      default: fin_eigen(rr, zf, zp, wvn, h, fg); break;
    }
    if (rr == 0.0) { fg.gr_re = 0.0; fg.gr_im = 0.0; }
  }
  fin_to_grn(fg, xpq, ypq, zf, zp, rr, !zero, flag, g);
}

// ---------------------------------------------------------------------------
// Infinite depth (Wu et al. 2017 Havelock approximation, InfGreen_Appr.f90)
// ---------------------------------------------------------------------------
HM_DEV double inf_poly(const double* c, int deg, double t) {
  double p = (c[deg - 1] + c[deg] * t) * t;
  for (int k = deg - 2; k >= 1; --k) p = (c[k] + p) * t;
  return c[0] + p;
}
HM_DEV double inf_j0(double x) {
  if (x <= 3.0) {
    const double y = (x / 3.0) * (x / 3.0);
    return +0.999999999 + (-2.249999879 + (+1.265623060 + (-0.316394552 + (+0.044460948 +
           (-0.003954479 + 0.000212950 * y) * y) * y) * y) * y) * y;
  }
  const double y = 3.0 / x, y2 = y * y;
  const double f0 = 0.79788454 + (-0.00553897 + (0.00099336 + (-0.00044346 + (0.00020445 - 0.00004959 * y2) * y2) * y2) * y2) * y2;
  const double th = x - 0.25 * kPi + (-0.04166592 + (0.00239399 + (-0.00073984 + (0.00031099 - 0.00007605 * y2) * y2) * y2) * y2) * y;
  return f0 * cos(th) / sqrt(x);
}
HM_DEV double inf_j1(double x) {
  if (x <= 3.0) {
    const double y = x / 3.0, y2 = y * y;
    return (+0.5 + (-0.562499992 + (+0.210937377 + (-0.039550040 + (+0.004447331 +
           (-0.000330547 + 0.000015525 * y2) * y2) * y2) * y2) * y2) * y2) * x;
  }
  const double y = 3.0 / x, y2 = y * y;
  const double f1 = 0.79788459 + (0.01662008 + (-0.00187002 + (0.00068519 + (-0.00029440 + 0.00006952 * y2) * y2) * y2) * y2) * y2;
  const double th = x - 0.75 * kPi + (0.12499895 + (-0.00605240 + (0.00135825 + (-0.00049616 + 0.00011531 * y2) * y2) * y2) * y2) * y;
  return f1 * cos(th) / sqrt(x);
}
HM_DEV double inf_y0(double x) {
  if (x <= 3.0) {
    const double q = (x / 3.0) * (x / 3.0);
    return (2.0 / kPi) * log(x / 2.0) * inf_j0(x) + 0.367466907 +
           (0.605593797 + (-0.743505078 + (0.253005481 + (-0.042619616 + (0.004285691 - 0.000250716 * q) * q) * q) * q) * q) * q;
  }
  const double q = 3.0 / x, q2 = q * q;
  const double f0 = 0.79788454 + (-0.00553897 + (0.00099336 + (-0.00044346 + (0.00020445 - 0.00004959 * q2) * q2) * q2) * q2) * q2;
  const double th = x - 0.25 * kPi + (-0.04166592 + (0.00239399 + (-0.00073984 + (0.00031099 - 0.00007605 * q2) * q2) * q2) * q2) * q;
  return f0 * sin(th) / sqrt(x);
}
HM_DEV double inf_y1(double x) {
  if (x <= 3.0) {
    const double q = x / 3.0, q2 = q * q;
    return (2.0 / kPi) * (log(x / 2.0) * inf_j1(x) - 1.0 / x) +
           (0.07373571 + (0.72276433 + (-0.43885620 + (0.10418264 + (-0.01340825 + 0.00094249 * q2) * q2) * q2) * q2) * q2) * q;
  }
  const double q = 3.0 / x, q2 = q * q;
  const double f1 = 0.79788459 + (0.01662008 + (-0.00187002 + (0.00068519 + (-0.00029440 + 0.00006952 * q2) * q2) * q2) * q2) * q2;
  const double th = x - 3.0 * kPi / 4.0 + (0.12499895 + (-0.00605240 + (0.00135825 + (-0.00049616 + 0.00011531 * q2) * q2) * q2) * q2) * q;
  return f1 * sin(th) / sqrt(x);
}
HM_DEV double inf_h0(double x) {
  if (x <= 3.0) {
    const double y = (x / 3.0) * (x / 3.0);
    return (+1.909859164 + (-1.909855001 + (+0.687514637 + (-0.126164557 + (+0.013828813 - 0.000876918 * y) * y) * y) * y) * y) * (x / 3.0);
  }
  const double y = (3.0 / x) * (3.0 / x);
  const double c1 = 2.0 * (0.99999906 + (4.77228920 + (3.85542044 + 0.32303607 * y) * y) * y);
  const double c2 = kPi * x * (1.0 + (4.88331068 + (4.28957333 + 0.52120508 * y) * y) * y);
  return c1 / c2 + inf_y0(x);
}
HM_DEV double inf_h1(double x) {
  if (x <= 3.0) {
    const double y = (x / 3.0) * (x / 3.0);
    return (+1.909859286 + (-1.145914713 + (+0.294656958 + (-0.042070508 + (+0.003785727 - 0.000207183 * y) * y) * y) * y) * y) * y;
  }
  const double y = (3.0 / x) * (3.0 / x);
  const double c1 = 2.0 * (1.00000004 + (3.92205313 + (2.64893033 + 0.27450895 * y) * y) * y);
  const double c2 = kPi * (1.0 + (3.81095112 + (2.26216956 + 0.10885141 * y) * y) * y);
  return c1 / c2 + inf_y1(x);
}

// Havelock function GF and its horizontal derivative GFh at (hh, vv) = (K R, K(z+zeta))
HM_DEV void havelock(double hh, double vv, double& gf_re, double& gf_im, double& gfh_re, double& gfh_im) {
  const double A[10] = {+1.21, -13.328, +215.896, -1763.96, +8418.94, -24314.21, +42002.57, -41592.9, 21859.0, -4838.6};
  const double B[10] = {+0.938, +5.373, -67.92, +796.534, -4780.77, +17137.74, -36618.81, +44894.06, -29030.24, +7671.22};
  const double C[8] = {+1.268, -9.747, +209.653, -1397.89, +5155.67, -9844.35, +9136.4, -3272.62};
  const double D[10] = {+0.632, -40.97, +667.16, -6072.07, +31127.39, -96293.05, +181856.75, -205690.43, +128170.2, -33744.6};
  const double dA[10] = {+2.948, -24.53, +249.69, -754.85, -1187.71, +16370.75, -48811.41, +68220.87, -46688.0, +12622.25};
  const double dB[10] = {+1.11, +2.894, -76.765, +1565.35, -11336.19, +44270.15, -97014.11, +118879.26, -76209.82, +19923.28};
  const double dC[6] = {+14.19, -148.24, +847.8, -2318.58, +3168.35, -1590.27};
  const double dd = sqrt(hh * hh + vv * vv);
  const double alpha = -vv / dd, beta = hh / dd, rho = dd / (1.0 + dd);
  const double ev = exp(vv), d3 = dd * dd * dd, om3 = (1.0 - rho) * (1.0 - rho) * (1.0 - rho);
  // L0 (Eq. 22), excluding -1/d
  double pp = log(0.5 * (dd - vv)) + 0.5772156649 - 2.0 * dd * dd;
  pp = ev * pp + dd * dd - vv;
  double rr = (1.0 - beta) * inf_poly(A, 9, rho) - beta * inf_poly(B, 9, rho);
  rr = rr - alpha * inf_poly(C, 7, rho) / (1.0 + 6.0 * alpha * rho * (1.0 - rho));
  rr = rr + beta * (1.0 - beta) * inf_poly(D, 9, rho);
  const double l0 = 2.0 * pp / (1.0 + d3) + 2.0 * (rho * om3 * rr);
  // Ls (Eq. 27)
  double ps = (beta + hh) / (dd - vv) - 2.0 * beta + 2.0 * ev * dd - hh;
  double qs = exp(-dd) * (1.0 - beta) * (1.0 + dd / (1.0 + d3));
  double r2 = beta * inf_poly(dA, 9, rho) - (1.0 - alpha) * inf_poly(dB, 9, rho);
  r2 = r2 + beta * (1.0 - beta) * rho * (1.0 - 2.0 * rho) * inf_poly(dC, 5, rho);
  const double ls = 2.0 * ps / (1.0 + d3) - 4.0 * qs + 2.0 * (rho * om3 * r2);
  // wave terms W, Wh
  const double h0 = inf_h0(hh), j0 = inf_j0(hh), h1 = inf_h1(hh), j1 = inf_j1(hh);
  gf_re = l0 + 2.0 * kPi * h0 * ev;
  gf_im = -2.0 * kPi * j0 * ev;
  gfh_re = ls + 2.0 * kPi * (2.0 / kPi - h1) * ev;
  gfh_im = 2.0 * kPi * j1 * ev;
}

// the closed-form imaginary parts of GF and GFh (the wave terms above)
HM_DEV void havelock_imag(double hh, double vv, double& gf_im, double& gfh_im) {
  const double ev = exp(vv);
  gf_im = -2.0 * kPi * inf_j0(hh) * ev;
  gfh_im = 2.0 * kPi * inf_j1(hh) * ev;
}

// the singular terms of Re GF and Re GFh near d = 0: the log(0.5 (d - vv)) of
// L0 and the (beta + hh) / (d - vv) of Ls, with their 2 / (1 + d^3) factors
HM_DEV void havelock_singular(double hh, double vv, double& s0, double& s1) {
  const double dd = sqrt(hh * hh + vv * vv), beta = hh / dd, d3 = dd * dd * dd;
  s0 = 2.0 * exp(vv) * log(0.5 * (dd - vv)) / (1.0 + d3);
  s1 = 2.0 * (beta + hh) / ((dd - vv) * (1.0 + d3));
}

// INFGREEN3D for a regular V (not the zero- or infinite-frequency limit) with
// the Havelock function GF, GFh given (the learned surrogate's path)
HM_DEV void infgreen3d_given(double xf, double xp, double yf, double yp, double zf, double zp, double v, int flag,
                             double gr, double gi, double ghr, double ghi, Grn4& g) {
  const double xpq = xf - xp, ypq = yf - yp, zpq = zp + zf;
  const double rr = horiz_dist(xf, xp, yf, yp);
  const double r = sqrt(rr * rr + (zp - zf) * (zp - zf)), r1 = sqrt(rr * rr + zpq * zpq);
  const double r13 = r1 * r1 * r1, r3 = r * r * r;
  for (int k = 0; k < 4; ++k) g.re[k] = g.im[k] = 0.0;
  if (rr > F32(1.e-6)) {
    g.re[1] = -(v * v) * ghr * xpq / rr; g.im[1] = -(v * v) * ghi * xpq / rr;
    g.re[2] = -(v * v) * ghr * ypq / rr; g.im[2] = -(v * v) * ghi * ypq / rr;
  }
  g.re[0] = -v * gr; g.im[0] = -v * gi;
  g.re[3] = -(v * v) * gr + 2.0 * v / r1; g.im[3] = -(v * v) * gi;
  g.re[0] += 1.0 / r1;
  g.re[1] += -xpq / r13;
  g.re[2] += -ypq / r13;
  g.re[3] += -zpq / r13;
  if (flag != 1) {
    g.re[0] += 1.0 / r;
    g.re[1] -= xpq / r3;
    g.re[2] -= ypq / r3;
    g.re[3] -= (zf - zp) / r3;
  }
}

HM_DEV void infgreen3d(double xf, double xp, double yf, double yp, double zf, double zp, double v,
                       int flag, Grn4& g) {
  const double xpq = xf - xp, ypq = yf - yp, zpq = zp + zf;
  const double rr = horiz_dist(xf, xp, yf, yp);
  const double r = sqrt(rr * rr + (zp - zf) * (zp - zf)), r1 = sqrt(rr * rr + zpq * zpq);
  const double r13 = r1 * r1 * r1, r3 = r * r * r;
  for (int k = 0; k < 4; ++k) g.re[k] = g.im[k] = 0.0;
  double sgn = 1.0;                       // +1: V=0 and finite V, -1: V=-1
  if (fabs(v) < F32(1.e-8)) {
  } else if (fabs(v + 1.0) < F32(1.e-8)) {
    sgn = -1.0;
  } else {
    double gr, gi, ghr, ghi;
    havelock(v * rr, v * zpq, gr, gi, ghr, ghi);
    if (rr > F32(1.e-6)) {
      g.re[1] = -(v * v) * ghr * xpq / rr; g.im[1] = -(v * v) * ghi * xpq / rr;
      g.re[2] = -(v * v) * ghr * ypq / rr; g.im[2] = -(v * v) * ghi * ypq / rr;
    }
    g.re[0] = -v * gr; g.im[0] = -v * gi;
    g.re[3] = -(v * v) * gr + 2.0 * v / r1; g.im[3] = -(v * v) * gi;
  }
  g.re[0] += sgn / r1;
  g.re[1] += -sgn * xpq / r13;
  g.re[2] += -sgn * ypq / r13;
  g.re[3] += -sgn * zpq / r13;
  if (flag != 1) {
    g.re[0] += 1.0 / r;
    g.re[1] -= xpq / r3;
    g.re[2] -= ypq / r3;
    g.re[3] -= (zf - zp) / r3;
  }
}

// ---------------------------------------------------------------------------
// Rankine panel integrals (Newman 1986), SingularIntgr.f90
// ---------------------------------------------------------------------------
struct Rk { double s, dx, dy, dz; };

// edge sum of one flat polygon in its local frame (ui, uj, uk); returns the
// potential and gradient contributions (st, sxt, syt, szt)
HM_DEV void rk_edges(int nv, const double* ksi, const double* eta, double x, double y, double z,
                     bool skip_short, double& st, double& sxt, double& syt, double& szt, double sgn) {
  st = sxt = syt = szt = 0.0;
  for (int i = 0; i < nv; ++i) {
    const int i1 = (i + 1 == nv) ? 0 : i + 1;
    const double dksi = ksi[i1] - ksi[i], deta = eta[i1] - eta[i];
    const double r1 = sqrt((ksi[i] - x) * (ksi[i] - x) + (eta[i] - y) * (eta[i] - y) + z * z);
    const double r2 = sqrt((ksi[i1] - x) * (ksi[i1] - x) + (eta[i1] - y) * (eta[i1] - y) + z * z);
    const double l = sqrt(dksi * dksi + deta * deta);
    if (skip_short && fabs(l) < 1.0e-8) continue;
    const double b1 = deta * ((ksi[i] - x) * (ksi[i] - x) + z * z) - dksi * (ksi[i] - x) * (eta[i] - y);
    const double a1 = r1 * z * dksi;
    const double b2 = deta * ((ksi[i1] - x) * (ksi[i1] - x) + z * z) - dksi * (ksi[i1] - x) * (eta[i1] - y);
    const double a2 = r2 * z * dksi;
    if (!(fabs(z) < 1.0e-6)) szt += atan2(b1 * a2 - b2 * a1, a1 * a2 + b1 * b2);
    if (fabs(r1 + r2 - l) < 1.0e-6) continue;
    const double lg = log((r1 + r2 + l) / (r1 + r2 - l));
    sxt -= deta / l * lg;
    syt += dksi / l * lg;
    st += sgn * (deta * (x - ksi[i]) - dksi * (y - eta[i])) / l * lg;
  }
}

// triangle (SGLINTBD_TRI :44); a quadrilateral is split into (1,2,3)+(1,3,4)
HM_DEV void rk_tri(const double* p0, const double* p1, const double* p2, const double* xf, Rk& o) {
  double p12[3], p13[3], po[3], po3[3], ui[3], uj[3], uk[3];
  for (int k = 0; k < 3; ++k) { p12[k] = p1[k] - p0[k]; p13[k] = p2[k] - p0[k]; }
  const double alfa = (p12[0] * p13[0] + p12[1] * p13[1] + p12[2] * p13[2]) /
                      (p12[0] * p12[0] + p12[1] * p12[1] + p12[2] * p12[2]);
  for (int k = 0; k < 3; ++k) { po[k] = p0[k] + alfa * p12[k]; po3[k] = p2[k] - po[k]; }
  const double l12 = sqrt(p12[0] * p12[0] + p12[1] * p12[1] + p12[2] * p12[2]);
  const double l3 = sqrt(po3[0] * po3[0] + po3[1] * po3[1] + po3[2] * po3[2]);
  for (int k = 0; k < 3; ++k) { ui[k] = p12[k] / l12; uj[k] = po3[k] / l3; }
  uk[0] = ui[1] * uj[2] - ui[2] * uj[1];
  uk[1] = ui[2] * uj[0] - ui[0] * uj[2];
  uk[2] = ui[0] * uj[1] - ui[1] * uj[0];
  double pof[3], d0[3], d1[3];
  for (int k = 0; k < 3; ++k) { pof[k] = xf[k] - po[k]; d0[k] = p0[k] - po[k]; d1[k] = p1[k] - po[k]; }
  const double x = ui[0] * pof[0] + ui[1] * pof[1] + ui[2] * pof[2];
  const double y = uj[0] * pof[0] + uj[1] * pof[1] + uj[2] * pof[2];
  const double z = uk[0] * pof[0] + uk[1] * pof[1] + uk[2] * pof[2];
  const double ksi[3] = {-copysign(1.0, alfa) * sqrt(d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2]), 0.0,
                         sqrt(d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2])};
  const double eta[3] = {0.0, l3, 0.0};
  double st, sxt, syt, szt;
  rk_edges(3, ksi, eta, x, y, z, false, st, sxt, syt, szt, 1.0);
  o.s += st - z * szt;
  o.dx += ui[0] * sxt + uj[0] * syt + uk[0] * szt;
  o.dy += ui[1] * sxt + uj[1] * syt + uk[1] * szt;
  o.dz += ui[2] * sxt + uj[2] * syt + uk[2] * szt;
}

// flat quadrilateral on its mean plane (SGLINTBD_QUAD :252)
HM_DEV void rk_quad(const double* v, const double* xf, Rk& o) {
  double po[3], pv[4][3], un[3], ui[3], uj[3], uk[3];
  for (int k = 0; k < 3; ++k) po[k] = (v[k] + v[3 + k] + v[6 + k] + v[9 + k]) / 4.0;
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 3; ++k) pv[i][k] = v[3 * i + k] - po[k];
  un[0] = pv[0][1] * pv[1][2] - pv[0][2] * pv[1][1];
  un[1] = pv[0][2] * pv[1][0] - pv[0][0] * pv[1][2];
  un[2] = pv[0][0] * pv[1][1] - pv[0][1] * pv[1][0];
  const double lp = sqrt(pv[0][0] * pv[0][0] + pv[0][1] * pv[0][1] + pv[0][2] * pv[0][2]);
  const double ln = sqrt(un[0] * un[0] + un[1] * un[1] + un[2] * un[2]);
  for (int k = 0; k < 3; ++k) { ui[k] = pv[0][k] / lp; uk[k] = un[k] / ln; }
  uj[0] = uk[1] * ui[2] - uk[2] * ui[1];
  uj[1] = uk[2] * ui[0] - uk[0] * ui[2];
  uj[2] = uk[0] * ui[1] - uk[1] * ui[0];
  double pof[3];
  for (int k = 0; k < 3; ++k) pof[k] = xf[k] - po[k];
  const double x = ui[0] * pof[0] + ui[1] * pof[1] + ui[2] * pof[2];
  const double y = uj[0] * pof[0] + uj[1] * pof[1] + uj[2] * pof[2];
  const double z = uk[0] * pof[0] + uk[1] * pof[1] + uk[2] * pof[2];
  double ksi[4], eta[4];
  for (int i = 0; i < 4; ++i) {
    const double dk = uk[0] * pv[i][0] + uk[1] * pv[i][1] + uk[2] * pv[i][2];
    double pn[3];
    for (int k = 0; k < 3; ++k) pn[k] = pv[i][k] - dk * uk[k];
    ksi[i] = ui[0] * pn[0] + ui[1] * pn[1] + ui[2] * pn[2];
    eta[i] = uj[0] * pn[0] + uj[1] * pn[1] + uj[2] * pn[2];
  }
  double st, sxt, syt, szt;
  rk_edges(4, ksi, eta, x, y, z, true, st, sxt, syt, szt, -1.0);
  o.s = st + z * szt;
  o.dx = -(ui[0] * sxt + uj[0] * syt + uk[0] * szt);
  o.dy = -(ui[1] * sxt + uj[1] * syt + uk[1] * szt);
  o.dz = -(ui[2] * sxt + uj[2] * syt + uk[2] * szt);
}

// ncn = 3 or 4 vertices in v[0..3*ncn)
HM_DEV void rankine(int ncn, const double* v, const double* xf, Rk& o) {
  if (ncn == 4) { rk_quad(v, xf, o); return; }
  o.s = o.dx = o.dy = o.dz = 0.0;
  rk_tri(v, v + 3, v + 6, xf, o);
}

}  // namespace hm
