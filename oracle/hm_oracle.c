/*
 * hm_oracle.c -- CPU restatement of the HAMS-MREL boundary-integral hot path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker. The tests, the smoke entry and
 * the CPU-baseline leg of the benchmark may call it; the product library never
 * does. Every function cites the reference file:line it restates. Single-
 * precision literals of the reference (Fortran REAL(4) constants such as 1.E-6,
 * -88.89 or the unsuffixed 1.0766115157 in JY01B) are reproduced as
 * float-rounded doubles, F32(...), because they change branch thresholds and
 * term counts.
 *
 * Parity status of each function (how it is pinned) is listed in the header
 * comment of tests/test_oracle.py.
 */
#include "hm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define F32(x) ((double)(float)(x))
static const double PI = 3.141592653589793;

/* ---- algorithmic operation counter ---------------------------------------- */
static double g_flops;
#define FL(n) (g_flops += (double)(n))
#define TR(n) (g_flops += (double)(n) * HMO_TRANS_FLOPS)
void hmo_flops_reset(void) { g_flops = 0.0; }
double hmo_flops_get(void) { return g_flops; }

/* =============================================================================
 * Special functions -- src/FinGrnExtSubs.f90 (Zhang & Jin, 1996)
 * ========================================================================== */

/* ERROR, FinGrnExtSubs.f90:62-95: erf(x) by series (|x|<3.5) or asymptotics. */
double hmo_error(double x)
{
    const double eps = 1.0e-15;
    double x2 = x * x, er, r, c0, err;
    FL(1);
    if (fabs(x) < 3.5) {
        er = 1.0; r = 1.0;
        for (int k = 1; k <= 50; k++) {
            r = r * x2 / (k + 0.5);
            er = er + r;
            FL(4);
            if (fabs(r) <= fabs(er) * eps) break;
        }
        c0 = 2.0 / sqrt(PI) * x * exp(-x2);
        err = c0 * er;
        FL(4); TR(1);
    } else {
        er = 1.0; r = 1.0;
        for (int k = 1; k <= 12; k++) {
            r = -r * (k - 0.5) / x2;
            er = er + r;
            FL(4);
        }
        c0 = exp(-x2) / (fabs(x) * sqrt(PI));
        err = 1.0 - c0 * er;
        if (x < 0.0) err = -err;
        FL(5); TR(1);
    }
    return err;
}

/* E1XA, FinGrnExtSubs.f90:198-211: E1(x) by rational approximations. */
double hmo_e1xa(double x)
{
    double e1, es1, es2;
    if (x == 0.0) {
        e1 = 1.0e300;
    } else if (x <= 1.0) {
        e1 = -log(x) + ((((1.07857e-3 * x - 9.76004e-3) * x + 5.519968e-2) * x - 0.24991055) * x
                        + 0.99999193) * x - 0.57721566;
        FL(12); TR(1);
    } else {
        es1 = (((x + 8.5733287401) * x + 18.059016973) * x + 8.6347608925) * x + 0.2677737343;
        es2 = (((x + 9.5733223454) * x + 25.6329561486) * x + 21.0996530827) * x + 3.9584969228;
        e1 = exp(-x) / x * es1 / es2;
        FL(19); TR(1);
    }
    return e1;
}

/* ENXA, FinGrnExtSubs.f90:135-155: En(x) by upward recurrence from E1. */
double hmo_enxa(int n, double x)
{
    double e0 = exp(-x) / x;
    double e1 = hmo_e1xa(x);
    double en = 0.0;
    TR(1); FL(1);
    if (n == 0) {
        en = e0;
    } else if (n == 1) {
        en = e1;
    } else {
        for (int k = 2; k <= n; k++) {
            en = (exp(-x) - x * e1) / (k - 1.0);
            e1 = en;
            TR(1); FL(4);
        }
    }
    return en;
}

/* EIX, FinGrnExtSubs.f90:254-281: Ei(x). */
double hmo_eix(double x)
{
    double ei, r;
    if (x == 0.0) {
        ei = -1.0e300;
    } else if (x <= 40.0) {
        ei = 1.0; r = 1.0;
        for (int k = 1; k <= 100; k++) {
            r = r * k * x / ((k + 1.0) * (k + 1.0));
            ei = ei + r;
            FL(6);
            if (fabs(r / ei) <= 1.0e-15) break;
        }
        ei = 0.5772156649015328 + log(x) + x * ei;
        FL(3); TR(1);
    } else {
        ei = 1.0; r = 1.0;
        for (int k = 1; k <= 20; k++) {
            r = r * k / x;
            ei = ei + r;
            FL(3);
        }
        ei = exp(x) / x * ei;
        FL(2); TR(1);
    }
    return ei;
}

/* Fortran integer power x**k (k >= 0) by repeated multiplication. */
static double ipow(double x, int k)
{
    double r = 1.0;
    for (int i = 0; i < k; i++) r *= x;
    return r;
}

/* IK01A, FinGrnExtSubs.f90:538-628: K0(x), K1(x) (the I0/I1 series are needed
 * for K1 and for the large-x K0). The reference reads an uninitialised WW on the
 * first pass of the K0 loop (:608); here it starts at 0, which never meets the
 * exit test on that pass. */
void hmo_ik01a(double x, double *bk0, double *bk1)
{
    static const double a[12] = {0.125, 7.03125e-2, 7.32421875e-2, 1.1215209960938e-1,
        2.2710800170898e-1, 5.7250142097473e-1, 1.7277275025845, 6.0740420012735,
        2.4380529699556e01, 1.1001714026925e02, 5.5133589612202e02, 3.0380905109224e03};
    static const double b[12] = {-0.375, -1.171875e-1, -1.025390625e-1, -1.4419555664063e-1,
        -2.7757644653320e-1, -6.7659258842468e-1, -1.9935317337513, -6.8839142681099,
        -2.7248827311269e01, -1.2159789187654e02, -6.0384407670507e02, -3.3022722944809e03};
    static const double a1[8] = {0.125, 0.2109375, 1.0986328125, 1.1775970458984e01,
        2.1461706161499e02, 5.9511522710323e03, 2.3347645606175e05, 1.2312234987631e07};
    const double el = 0.5772156649015329;
    double x2 = x * x, bi0, bi1, r, k0v, k1v;
    FL(1);
    if (x == 0.0) {
        *bk0 = 1.0e300;
        *bk1 = 1.0e300;
        return;
    } else if (x <= 18.0) {
        bi0 = 1.0; r = 1.0;
        for (int k = 1; k <= 50; k++) {
            r = 0.25 * r * x2 / (k * k);
            bi0 = bi0 + r;
            FL(6);
            if (fabs(r / bi0) < 1.0e-15) break;
        }
        bi1 = 1.0; r = 1.0;
        for (int k = 1; k <= 50; k++) {
            r = 0.25 * r * x2 / (k * (k + 1));
            bi1 = bi1 + r;
            FL(6);
            if (fabs(r / bi1) < 1.0e-15) break;
        }
        bi1 = 0.5 * x * bi1;
        FL(2);
    } else {
        int k0 = 12;
        if (x >= 35.0) k0 = 9;
        if (x >= 50.0) k0 = 7;
        double ca = exp(x) / sqrt(2.0 * PI * x);
        double xr = 1.0 / x;
        bi0 = 1.0;
        for (int k = 1; k <= k0; k++) { bi0 = bi0 + a[k - 1] * ipow(xr, k); FL(2 + k); }
        bi0 = ca * bi0;
        bi1 = 1.0;
        for (int k = 1; k <= k0; k++) { bi1 = bi1 + b[k - 1] * ipow(xr, k); FL(2 + k); }
        bi1 = ca * bi1;
        FL(7); TR(1);
    }
    if (x <= 9.0) {
        double ct = -(log(x / 2.0) + el);
        double w0 = 0.0, ww = 0.0;
        k0v = 0.0; r = 1.0;
        TR(1); FL(3);
        for (int k = 1; k <= 50; k++) {
            w0 = w0 + 1.0 / k;
            r = 0.25 * r / (k * k) * x2;
            k0v = k0v + r * (w0 + ct);
            FL(10);
            if (fabs((k0v - ww) / k0v) < 1.0e-15) break;
            ww = k0v;
        }
        k0v = k0v + ct;
        FL(1);
    } else {
        double cb = 0.5 / x, xr2 = 1.0 / x2;
        k0v = 1.0;
        for (int k = 1; k <= 8; k++) { k0v = k0v + a1[k - 1] * ipow(xr2, k); FL(2 + k); }
        k0v = cb * k0v / bi0;
        FL(4);
    }
    k1v = (1.0 / x - bi1 * k0v) / bi0;
    FL(4);
    *bk0 = k0v;
    *bk1 = k1v;
}

/* JY01B, FinGrnExtSubs.f90:663-705 (and its J-only / Y-only twins JY01BJ :741,
 * JY01BY0 :811, JY01BY1 :875, which evaluate the same polynomials). Note the
 * REAL(4) literal 1.0766115157 in the Y0 polynomial (:681). */
void hmo_jy01b(double x, double *bj0, double *bj1, double *by0, double *by1)
{
    double t, t2, j0, j1, y0, y1;
    if (x == 0.0) {
        *bj0 = 1.0; *bj1 = 0.0; *by0 = -1.0e300; *by1 = -1.0e300;
        return;
    } else if (x <= 4.0) {
        t = x / 4.0;
        t2 = t * t;
        j0 = ((((((-.5014415e-3 * t2 + .76771853e-2) * t2 - .0709253492) * t2 + .4443584263)
              * t2 - 1.7777560599) * t2 + 3.9999973021) * t2 - 3.9999998721) * t2 + 1.0;
        j1 = t * (((((((-.1289769e-3 * t2 + .22069155e-2) * t2 - .0236616773) * t2 + .1777582922)
              * t2 - .8888839649) * t2 + 2.6666660544) * t2 - 3.9999999710) * t2 + 1.9999999998);
        y0 = (((((((-.567433e-4 * t2 + .859977e-3) * t2 - .94855882e-2) * t2 + .0772975809) * t2
              - .4261737419) * t2 + 1.4216421221) * t2 - 2.3498519931) * t2 + F32(1.0766115157)) * t2
             + .3674669052;
        y0 = 2.0 / PI * log(x / 2.0) * j0 + y0;
        y1 = ((((((((.6535773e-3 * t2 - .0108175626) * t2 + .107657606) * t2 - .7268945577) * t2
              + 3.1261399273) * t2 - 7.3980241381) * t2 + 6.8529236342) * t2 + .3932562018) * t2
              - .6366197726) / x;
        y1 = 2.0 / PI * log(x / 2.0) * j1 + y1;
        FL(80); TR(2);
    } else {
        t = 4.0 / x;
        t2 = t * t;
        double a0 = sqrt(2.0 / (PI * x));
        double p0 = ((((-.9285e-5 * t2 + .43506e-4) * t2 - .122226e-3) * t2 + .434725e-3) * t2
                     - .4394275e-2) * t2 + .999999997;
        double q0 = t * (((((.8099e-5 * t2 - .35614e-4) * t2 + .85844e-4) * t2 - .218024e-3) * t2
                     + .1144106e-2) * t2 - .031249995);
        double ta0 = x - .25 * PI;
        j0 = a0 * (p0 * cos(ta0) - q0 * sin(ta0));
        y0 = a0 * (p0 * sin(ta0) + q0 * cos(ta0));
        double p1 = ((((.10632e-4 * t2 - .50363e-4) * t2 + .145575e-3) * t2 - .559487e-3) * t2
                     + .7323931e-2) * t2 + 1.000000004;
        double q1 = t * (((((-.9173e-5 * t2 + .40658e-4) * t2 - .99941e-4) * t2 + .266891e-3) * t2
                     - .1601836e-2) * t2 + .093749994);
        double ta1 = x - .75 * PI;
        j1 = a0 * (p1 * cos(ta1) - q1 * sin(ta1));
        y1 = a0 * (p1 * sin(ta1) + q1 * cos(ta1));
        FL(75); TR(4);
    }
    *bj0 = j0; *bj1 = j1; *by0 = y0; *by1 = y1;
}

/* Limes, FinGrnExtSubs.f90:370-503: Wynn epsilon-algorithm limit of the
 * partial sums s[0..n] (Mishonov & Penev 2000). s is overwritten, exactly as the
 * reference's INOUT argument. Exact floating-point equality tests are kept. */
double hmo_limes(int n, double *s)
{
    double a[256];
    double rl = s[n], amax = 0.0;
    int k = 1;
    for (int i = 0; i <= n; i++) a[i] = 0.0;
    for (;;) {
        if (n - 2 * k + 1 < 0) break;
        for (int i = 0; i <= n - 2 * k + 1; i++) {
            if (s[i + 1] != s[i]) a[i] = a[i + 1] + 1.0 / (s[i + 1] - s[i]);
            else a[i] = a[i + 1];
            FL(3);
        }
        if (n - 2 * k < 0) break;
        for (int i = 0; i <= n - 2 * k; i++) {
            if (a[i + 1] != a[i]) s[i] = s[i + 1] + 1.0 / (a[i + 1] - a[i]);
            else s[i] = s[i + 1];
            FL(3);
            if (fabs(a[i]) > amax) {
                amax = fabs(a[i]);
                rl = s[i];
                if (s[i + 1] == s[i]) return rl;
            }
        }
        k = k + 1;
    }
    return rl;
}

/* =============================================================================
 * Dispersion relation -- src/WavDynSubs.f90:40-127 (Newman 1990, third-order
 * iteration). g_disp is the gravity used inside the solver (9.807 in the
 * reference, :58).
 * ========================================================================== */
void hmo_dispersion(double *wvn, int nrt, double w, double h, double g_disp)
{
    double x = w * w * h / g_disp, y, t, u, trial, dnm, fun, dfun, d2fun, exx;
    if (x > 0.0 && x <= 2.0) {
        y = sqrt(x) * (0.9994 + 0.1701 * x + F32(0.0305) * x * x);
    } else {
        t = x * exp(-2.0 * x);
        y = x + 2.0 * t - 6.0 * t * t;
    }
    if (fabs(y - x) < F32(1.e-10)) {
        wvn[0] = x / h;
    } else {
        exx = 1.0;
        while (exx > 1.0e-10) {
            trial = y;
            dnm = trial * trial - x * x;
            fun = log((trial + x) / (trial - x)) / 2.0 - trial;
            dfun = -x / dnm - 1.0;
            d2fun = 2.0 * x * trial / (dnm * dnm);
            y = trial - fun / dfun * (1.0 + (fun / dfun) * (d2fun / dfun) / 2.0);
            exx = fabs(y - trial);
        }
        wvn[0] = y / h;
    }
    u = 3.0 * x / (7.0 + 3.0 * x);
    t = 0.0159 + 0.1032 * u + 4.3152 * u * u - 2.8768 * u * u * u;
    for (int i = 2; i <= nrt; i++) {
        exx = 1.0;
        while (exx > 1.0e-10) {
            trial = t;
            y = (i - 1) * PI - trial;
            dnm = y * y + x * x;
            fun = atan2(x, y) - trial;
            dfun = x / dnm - 1.0;
            d2fun = 2.0 * x * trial / (dnm * dnm);
            t = trial - fun / dfun * (1.0 + (fun / dfun) * (d2fun / dfun) / 2.0);
            exx = fabs(t - trial);
        }
        y = (i - 1) * PI - t;
        wvn[i - 1] = y / h;
        t = t - PI * x / (x * x + PI * i * (PI * (i - 1) - t));
    }
}

/* =============================================================================
 * Finite-depth Green function -- src/FinGreen3D.f90 (Liu et al., FinGreen3D)
 * ========================================================================== */

static void singular_strip(double r, double zf, double zp, double complex g[3])
{
    /* the common tail of EIGEN/EIGENE/PIDCOCK (FinGreen3D.f90:349-358): remove
     * 1/r + 1/r1 and their derivatives for the panel-integration scheme */
    double sr = sqrt(r * r + (zf - zp) * (zf - zp));
    double sr1 = sqrt(r * r + (zf + zp) * (zf + zp));
    double dsrr = r / sr, dsrz = (zf - zp) / sr, dsr1r = r / sr1, dsr1z = (zf + zp) / sr1;
    g[0] = g[0] - (1.0 / sr + 1.0 / sr1);
    g[1] = g[1] + (dsrr / (sr * sr) + dsr1r / (sr1 * sr1));
    g[2] = g[2] + (dsrz / (sr * sr) + dsr1z / (sr1 * sr1));
    FL(30);
}

/* EIGEN, FinGreen3D.f90:281-362: eigenfunction series, R/h >= 0.5. */
static void eigen(double r, double zf, double zp, double wk, const double *wvn, double h,
                  double complex g[3])
{
    const int nt = 10;
    double zfh = zf + h, zph = zp + h, j0, j1, y0, y1;
    double nm0 = h / 2.0 * (1.0 + sinh(2.0 * wk * h) / (2.0 * wk * h));
    double p0 = PI / nm0 * cosh(wk * zfh) * cosh(wk * zph);
    double dp0 = PI / nm0 * sinh(wk * zfh) * cosh(wk * zph);
    FL(16); TR(5);
    hmo_jy01b(wk * r, &j0, &j1, &y0, &y1);
    g[0] = p0 * (-y0 + I * j0);
    g[1] = -wk * p0 * (-y1 + I * j1);
    g[2] = wk * dp0 * (-y0 + I * j0);
    FL(8);
    for (int i = 2; i <= nt; i++) {
        double km = wvn[i - 1], k0, k1;
        double twokmh = 2.0 * km * h;
        double pm = 4.0 / (1.0 + sin(twokmh) / twokmh) / h;
        hmo_ik01a(km * r, &k0, &k1);
        double kzf = km * zfh, kzp = km * zph;
        double ckzf = cos(kzf), ckzp = cos(kzp), skzf = sin(kzf);
        double rm1 = pm * k0 * ckzf * ckzp;
        double rm2 = -pm * km * k1 * ckzf * ckzp;
        double rm3 = -pm * km * k0 * skzf * ckzp;
        g[0] += rm1; g[1] += rm2; g[2] += rm3;
        FL(22); TR(4);
        if (fabs(rm1) < F32(1.e-6) && fabs(rm2) < F32(1.e-6) && fabs(rm3) < F32(1.e-6)) break;
    }
    singular_strip(r, zf, zp, g);
}

/* EIGENE, FinGreen3D.f90:413-506: eigenfunction series with Wynn-epsilon
 * acceleration, 0.05 <= R/h < 0.5. NT = INT(-88.89*R/h + 54.45) in REAL(4)
 * constants (:433). */
static void eigene(double r, double zf, double zp, double wk, const double *wvn, int nk,
                   double h, double complex g[3])
{
    double rhp = fabs(r / h);
    int nt = (int)(F32(-88.89) * rhp + F32(54.45));
    double zfh = zf + h, zph = zp + h, j0, j1, y0, y1;
    double gs[256], gr[256], gz[256];
    double nm0 = h / 2.0 * (1.0 + sinh(2.0 * wk * h) / (2.0 * wk * h));
    double p0 = PI / nm0 * cosh(wk * zfh) * cosh(wk * zph);
    double dp0 = PI / nm0 * sinh(wk * zfh) * cosh(wk * zph);
    (void)nk;
    FL(20); TR(5);
    hmo_jy01b(wk * r, &j0, &j1, &y0, &y1);
    g[0] = p0 * (-y0 + I * j0);
    g[1] = -wk * p0 * (-y1 + I * j1);
    g[2] = wk * dp0 * (-y0 + I * j0);
    FL(8);
    gs[0] = 0.0; gr[0] = 0.0; gz[0] = 0.0;
    for (int i = 2; i <= nt; i++) {
        double km = wvn[i - 1], k0, k1;
        double twokmh = 2.0 * km * h;
        double pm = 4.0 / (1.0 + sin(twokmh) / twokmh) / h;
        hmo_ik01a(km * r, &k0, &k1);
        double kzf = km * zfh, kzp = km * zph;
        double ckzf = cos(kzf), ckzp = cos(kzp), skzf = sin(kzf);
        gs[i - 1] = gs[i - 2] + pm * k0 * ckzf * ckzp;
        gr[i - 1] = gr[i - 2] + (-pm * km * k1 * ckzf * ckzp);
        gz[i - 1] = gz[i - 2] + (-pm * km * k0 * skzf * ckzp);
        FL(22); TR(4);
    }
    g[0] += hmo_limes(nt - 1, gs);
    g[1] += hmo_limes(nt - 1, gr);
    g[2] += hmo_limes(nt - 1, gz);
    singular_strip(r, zf, zp, g);
}

/* Test hook: the EIGENE series terms t[0..nt-1] (t[0] = 0) whose running sums
 * eigene() feeds to Limes; lets a test form the partial sums in another order. */
int hmo_eigene_terms(double r, double zf, double zp, const double *wvn, double h, double *t1, double *t2,
                     double *t3)
{
    double rhp = fabs(r / h);
    int nt = (int)(F32(-88.89) * rhp + F32(54.45));
    double zfh = zf + h, zph = zp + h;
    t1[0] = t2[0] = t3[0] = 0.0;
    for (int i = 2; i <= nt; i++) {
        double km = wvn[i - 1], k0, k1;
        double twokmh = 2.0 * km * h;
        double pm = 4.0 / (1.0 + sin(twokmh) / twokmh) / h;
        hmo_ik01a(km * r, &k0, &k1);
        double kzf = km * zfh, kzp = km * zph;
        double ckzf = cos(kzf), ckzp = cos(kzp), skzf = sin(kzf);
        t1[i - 1] = pm * k0 * ckzf * ckzp;
        t2[i - 1] = -pm * km * k1 * ckzf * ckzp;
        t3[i - 1] = -pm * km * k0 * skzf * ckzp;
    }
    return nt;
}

static const double AMN[25] = {   /* AMN(0:4,0:4), column-major DATA (FinGreen3D.f90:949) */
    0.0, -0.02525711, 0.00086546, -0.00004063, 0.00000193,
    0.05051418, -0.00692380, 0.00073292, -0.00006636, 0.00000398,
    0.00230838, -0.00097875, 0.00020597, -0.00003333, 0.00000524,
    0.00012934, -0.00010879, 0.00003965, -0.00000891, 0.0,
    0.00000913, -0.00001270, 0.00000466, 0.0, 0.0};
#define AMN_(m, n) AMN[(m) + 5 * (n)]

static double rs(double r, double z) { FL(4); return sqrt(r * r + z * z); }

/* IMGS, FinGreen3D.f90:941-971: Chebyshev fit of the Rankine image sum. */
static double imgs(double r, double zf, double zp, double h)
{
    double t1 = r / h, t2 = (zf - zp) / h, t3 = (zf + zp + 2.0 * h) / h, s1 = 0.0, s2 = 0.0;
    for (int l = -1; l <= 1; l++) s1 = s1 + 1.0 / rs(t1, t2 + 2.0 * l) + 1.0 / rs(t1, t3 + 2.0 * l);
    for (int m = 0; m <= 4; m++)
        for (int n = 0; n <= 4; n++)
            s2 = s2 + AMN_(m, n) * pow(t1, 2.0 * m) * (pow(t2, 2.0 * n) + pow(t3, 2.0 * n));
    FL(8 + 15 + 25 * 6); TR(75);
    return (s1 - 2.0 + s2) / h;
}

/* DGSR, FinGreen3D.f90:1006-1036: R-derivative of IMGS. */
static double dgsr(double r, double zf, double zp, double h)
{
    double t1 = r / h, t2 = (zf - zp) / h, t3 = (zf + zp + 2.0 * h) / h, s1 = 0.0, s2 = 0.0;
    for (int l = -1; l <= 1; l++) {
        double a = rs(t1, t2 + 2.0 * l), b = rs(t1, t3 + 2.0 * l);
        s1 = s1 + t1 / (a * a * a) + t1 / (b * b * b);
    }
    for (int m = 1; m <= 4; m++)
        for (int n = 0; n <= 4; n++)
            s2 = s2 + 2.0 * m * AMN_(m, n) * pow(t1, 2.0 * m - 1.0) * (pow(t2, 2.0 * n) + pow(t3, 2.0 * n));
    FL(8 + 30 + 20 * 9); TR(60);
    return (-s1 + s2) / (h * h);
}

/* DGSZ, FinGreen3D.f90:1085-1114: z-derivative of IMGS. */
static double dgsz(double r, double zf, double zp, double h)
{
    double t1 = r / h, t2 = (zf - zp) / h, t3 = (zf + zp + 2.0 * h) / h, s1 = 0.0, s2 = 0.0;
    for (int l = -1; l <= 1; l++) {
        double a = rs(t1, t2 + 2.0 * l), b = rs(t1, t3 + 2.0 * l);
        s1 = s1 + (t2 + 2.0 * l) / (a * a * a) + (t3 + 2.0 * l) / (b * b * b);
    }
    for (int m = 0; m <= 4; m++)
        for (int n = 1; n <= 4; n++)
            s2 = s2 + 2.0 * n * AMN_(m, n) * pow(t1, 2.0 * m) * (pow(t2, 2.0 * n - 1.0) + pow(t3, 2.0 * n - 1.0));
    FL(8 + 36 + 20 * 10); TR(60);
    return (-s1 + s2) / (h * h);
}

/* PIDCOCK, FinGreen3D.f90:558-660: series accelerated by subtracting an
 * asymptotic series, Wynn-epsilon limit, plus the Chebyshev image sums;
 * 0.0005 <= R/h < 0.05. NT = INT(-1010.10*R/h + 100.50), REAL(4) constants. */
static void pidcock(double r, double zf, double zp, double v, double wk, const double *wvn,
                    double h, double complex g[3])
{
    const double ga = 0.5772156649015328;
    double j0, j1, y0, y1, gs[256], gr[256], gz[256];
    double nm0 = h / 2.0 * (1.0 + sinh(2.0 * wk * h) / (2.0 * wk * h));
    double p0 = PI / nm0 * cosh(wk * (zf + h)) * cosh(wk * (zp + h));
    double dp0 = wk * PI / nm0 * sinh(wk * (zf + h)) * cosh(wk * (zp + h));
    FL(22); TR(5);
    hmo_jy01b(wk * r, &j0, &j1, &y0, &y1);
    g[0] = p0 * (-y0 + I * j0) + 2.0 / h * (ga + log(r / 4.0 / h));
    g[1] = -wk * p0 * (-y1 + I * j1) + 2.0 / (r * h);
    g[2] = dp0 * (-y0 + I * j0);
    FL(16); TR(1);
    double rhp = fabs(r / h);
    int nt = (int)(F32(-1010.10) * rhp + F32(100.50));
    gs[0] = 0.0; gr[0] = 0.0; gz[0] = 0.0;
    for (int i = 2; i <= nt; i++) {
        double mp = (i - 1) * PI, umt = mp / h, um = wvn[i - 1], x = um * r, k0, k1, k0t, k1t;
        hmo_ik01a(x, &k0, &k1);
        hmo_ik01a(umt * r, &k0t, &k1t);
        double ac = (um * um + v * v) / ((um * um + v * v) * h - v);
        double am = ac * cos(um * (zf + h)) * cos(um * (zp + h));
        double amt = cos(umt * (zf + h)) * cos(umt * (zp + h)) / h;
        double dam = -um * ac * sin(um * (zf + h)) * cos(um * (zp + h));
        double damt = -umt * sin(umt * (zf + h)) * cos(umt * (zp + h)) / h;
        gs[i - 1] = gs[i - 2] + (am * k0 - amt * k0t);
        gr[i - 1] = gr[i - 2] + (-um * am * k1 + umt * amt * k1t);
        gz[i - 1] = gz[i - 2] + (dam * k0 - damt * k0t);
        FL(48); TR(8);
    }
    double rl1 = hmo_limes(nt - 1, gs), rl2 = hmo_limes(nt - 1, gr), rl3 = hmo_limes(nt - 1, gz);
    g[0] = g[0] + 4.0 * rl1 + imgs(r, zf, zp, h);
    g[1] = g[1] + 4.0 * rl2 + dgsr(r, zf, zp, h);
    g[2] = g[2] + 4.0 * rl3 + dgsz(r, zf, zp, h);
    FL(9);
    singular_strip(r, zf, zp, g);
}

/* ---- LINTON helpers -------------------------------------------------------- */

/* ERFCC / DERFCC, FinGreen3D.f90:2301-2320 */
static double erfcc(double x) { FL(1); return 1.0 - hmo_error(x); }
static double derfcc(double x) { FL(4); TR(2); return -2.0 / sqrt(PI) * exp(-pow(x, 2.0)); }

/* CHI / DCHI, FinGreen3D.f90:2253-2285 */
static void chi(double zf, double zp, double h, double val[4])
{
    val[0] = -zp - zf;
    val[1] = 2.0 * h - zp + zf;
    val[2] = 2.0 * h + zp - zf;
    val[3] = 4.0 * h + zp + zf;
    FL(10);
}
static const double DVAL[4] = {-1.0, 1.0, -1.0, 1.0};

/* RCHI, FinGreen3D.f90:2216-2240 */
static void rchi(double r, double zf, double zp, double h, double ro[4], double dror[4],
                 double droz[4])
{
    double val[4];
    chi(zf, zp, h, val);
    for (int i = 0; i < 4; i++) {
        double sgn = ((i + 1) % 2) ? -1.0 : 1.0;   /* (-1)**I, I = 1..4 */
        ro[i] = sqrt(r * r + val[i] * val[i]);
        dror[i] = r / sqrt(r * r + val[i] * val[i]);
        droz[i] = val[i] * sgn / sqrt(r * r + val[i] * val[i]);
        FL(15);
    }
}

/* NM, FinGreen3D.f90:2289-2297: normalisation of the complex eigenfunction. */
static double complex nmf(double h, double complex wvn)
{
    FL(12); TR(1);
    return h / 2.0 * (1.0 + csin(2.0 * wvn * h) / (2.0 * wvn * h));
}

/* F1 / F2, FinGreen3D.f90:1800-1865: integrands of the evanescent coefficients. */
static double f1(double r, double wvn, int no, double t)
{
    if (fabs(t) < 1.0e-8) return 0.0;
    FL(10); TR(1);
    if (no == 1) return exp(-r * r / 4.0 / t + wvn * wvn * t) / 4.0 / PI / t;
    return exp(-r * r / 4.0 / t - wvn * wvn * t) / 4.0 / PI / t;
}
static double f2(double r, double wvn, int no, double t)
{
    if (fabs(t) < 1.0e-8) return 0.0;
    FL(12); TR(1);
    if (no == 1) return -r * exp(-r * r / 4.0 / t + wvn * wvn * t) / 8.0 / PI / (t * t);
    return -r * exp(-r * r / 4.0 / t - wvn * wvn * t) / 8.0 / PI / (t * t);
}

/* G1 / G2 / G3, FinGreen3D.f90:1870-1990: integrands of the Linton
 * free-surface integral. */
static double g1f(double v, double r, double zf, double zp, double h, double t)
{
    double val[4], sum = 0.0;
    if (fabs(t) < 1.0e-8) return 0.0;
    chi(zf, zp, h, val);
    for (int i = 0; i < 4; i++) { sum = sum + exp(-v * val[i]) * erfcc(val[i] / 2.0 / t - v * t); FL(7); TR(1); }
    FL(12); TR(1);
    return exp(v * v * t * t - r * r / 4.0 / (t * t)) * sum / t;
}
static double g2f(double v, double r, double zf, double zp, double h, double t)
{
    double val[4], sum = 0.0;
    if (fabs(t) < 1.0e-8) return 0.0;
    chi(zf, zp, h, val);
    for (int i = 0; i < 4; i++) { sum = sum + exp(-v * val[i]) * erfcc(val[i] / 2.0 / t - v * t); FL(7); TR(1); }
    FL(16); TR(1);
    return r / 2.0 / (t * t * t) * exp(v * v * t * t - r * r / 4.0 / (t * t)) * sum;
}
static double g3f(double v, double r, double zf, double zp, double h, double t)
{
    double val[4], sum = 0.0;
    if (fabs(t) < 1.0e-8) return 0.0;
    chi(zf, zp, h, val);
    for (int i = 0; i < 4; i++) {
        double arg = val[i] / 2.0 / t - v * t;
        sum = sum - v * DVAL[i] * exp(-v * val[i]) * erfcc(arg)
                  + exp(-v * val[i]) * derfcc(arg) * DVAL[i] / 2.0 / t;
        FL(16); TR(2);
    }
    FL(12); TR(1);
    return exp(v * v * t * t - r * r / 4.0 / (t * t)) * sum / t;
}

static double funf(double v, double r, double zf, double zp, double h, double t, int tag)
{
    if (tag == 1) return g1f(v, r, zf, zp, h, t);
    if (tag == 2) return g2f(v, r, zf, zp, h, t);
    return g3f(v, r, zf, zp, h, t);
}
static double fun(double r, double wvn, int no, double t, int tag)
{
    return tag == 1 ? f1(r, wvn, no, t) : f2(r, wvn, no, t);
}

/* 7-point Gauss / 15-point Kronrod nodes and weights (FinGreen3D.f90:2403-2431) */
static const double XGQ[7] = {-0.949107912342759, -0.741531185599394, -0.405845151377397, 0.0,
                              0.405845151377397, 0.741531185599394, 0.949107912342759};
static const double WGQ[7] = {0.129484966168870, 0.279705391489277, 0.381830050505119,
                              0.417959183673469, 0.381830050505119, 0.279705391489277,
                              0.129484966168870};
static const double XKQ[15] = {-0.991455371120813, -0.949107912342759, -0.864864423359769,
                               -0.741531185599394, -0.586087235467691, -0.405845151377397,
                               -0.207784955007898, 0.0, 0.207784955007898, 0.405845151377397,
                               0.586087235467691, 0.741531185599394, 0.864864423359769,
                               0.949107912342759, 0.991455371120813};
static const double WKQ[15] = {0.022935322010529, 0.063092092629979, 0.104790010322250,
                               0.140653259715525, 0.169004726639267, 0.190350578064785,
                               0.204432940075298, 0.209482141084728, 0.204432940075298,
                               0.190350578064785, 0.169004726639267, 0.140653259715525,
                               0.104790010322250, 0.063092092629979, 0.022935322010529};

/* GK_INTEGF + recursive AQUADF, FinGreen3D.f90:2387-2532 (and the twins
 * GK_INTEG/AQUAD :2548-2745): adaptive bisection until the G7-K15 error
 * estimate (200|K-G|)^1.5 is below eps or the interval is shorter than eps. */
typedef struct { double v, r, zf, zp, h, wvn; int no, tag, kind; } quad_ctx;

static double quad_f(const quad_ctx *c, double t)
{
    return c->kind == 0 ? funf(c->v, c->r, c->zf, c->zp, c->h, t, c->tag)
                        : fun(c->r, c->wvn, c->no, t, c->tag);
}

static double aquad_rec(const quad_ctx *c, double a, double b, double eps)
{
    double stf[7], gint = 0.0, gk = 0.0;
    for (int i = 0; i < 7; i++) {
        stf[i] = quad_f(c, 1 / 2.0 * (b + a + (b - a) * XGQ[i]));
        gint = gint + WGQ[i] * (b - a) / 2.0 * stf[i];
        FL(9);
    }
    for (int i = 1; i <= 15; i++) {
        if (i % 2 == 0) gk = gk + WKQ[i - 1] * (b - a) / 2.0 * stf[i / 2 - 1];
        else gk = gk + WKQ[i - 1] * (b - a) / 2.0 * quad_f(c, 1 / 2.0 * (b + a + (b - a) * XKQ[i - 1]));
        FL(9);
    }
    double err = pow(200.0 * fabs(gk - gint), 1.5);
    FL(3); TR(1);
    if (err > eps && fabs(a - b) > eps)
        return aquad_rec(c, a, (a + b) / 2.0, eps) + aquad_rec(c, (a + b) / 2.0, b, eps);
    return gk;
}

static double aquadf(double v, double r, double zf, double zp, double h, double a, double b,
                     double eps, int tag)
{
    quad_ctx c = {v, r, zf, zp, h, 0.0, 0, tag, 0};
    return aquad_rec(&c, a, b, eps);
}
static double aquad(double r, double wvn, int no, double a, double b, double eps, int tag)
{
    quad_ctx c = {0.0, r, 0.0, 0.0, 0.0, wvn, no, tag, 1};
    return aquad_rec(&c, a, b, eps);
}

/* SA0 / DSA0 / SAM / DSAM, FinGreen3D.f90:1389-1694: small-R series of the
 * Linton expansion coefficients. */
static double sa0(double r, double pa, double wk)
{
    const double ga = 0.5772156649015328, eps = F32(1.e-6);
    int n = 1;
    double rm = 100.0, s = -ga - 2.0 * log(wk * pa / 2.0), sgn = -1.0, facn = 1.0;
    TR(1); FL(4);
    while (fabs(rm) > eps) {
        double en = hmo_enxa(n + 1, (r / pa) * (r / pa)), psi = 0.0;
        for (int p = 1; p <= n; p++) psi = psi + 1.0 / p;
        rm = sgn * pow(r / pa, 2.0 * n) / facn / n - pow(wk * pa / 2.0, 2.0 * n) / facn * en
             - 2.0 * sgn / (facn * facn) * pow(wk * r / 2.0, 2.0 * n) * (log(wk * r / 2.0) + ga - psi);
        s = s + rm;
        n = n + 1;
        sgn = -sgn;
        facn = facn * n;
        FL(30 + 2 * n); TR(4);
    }
    return s / 4.0 / PI;
}
static double dsa0(double r, double pa, double wk)
{
    const double ga = 0.5772156649015328, eps = F32(1.e-6);
    int n = 1;
    double rm = 100.0, s = 0.0, sgn = -1.0, facn = 1.0;
    while (fabs(rm) > eps) {
        double en = hmo_enxa(n, (r / pa) * (r / pa)), psi = 0.0;
        for (int p = 1; p <= n; p++) psi = psi + 1.0 / p;
        rm = 2.0 * sgn * pow(r / pa, 2.0 * n - 1.0) / facn / pa
             + pow(wk * pa / 2.0, 2.0 * n) / facn * en * 2.0 * r / (pa * pa)
             - 2.0 * sgn * n * wk / (facn * facn) * pow(wk * r / 2.0, 2.0 * n - 1.0)
                   * (log(wk * r / 2.0) + ga - psi)
             - 2.0 * sgn / (facn * facn) * pow(wk * r / 2.0, 2.0 * n) / r;
        s = s + rm;
        n = n + 1;
        sgn = -sgn;
        facn = facn * n;
        FL(45 + 2 * n); TR(5);
    }
    return s / 4.0 / PI;
}
static double sam(double r, double pa, double wvn)
{
    const double eps = F32(1.e-6);
    int n = 0;
    double rm = 100.0, s = 0.0, sgn = 1.0, facn = 1.0;
    while (fabs(rm) > eps) {
        double en = hmo_enxa(n + 1, (wvn * pa / 2.0) * (wvn * pa / 2.0));
        rm = sgn * pow(r / pa, 2.0 * n) / facn * en;
        s = s + rm;
        n = n + 1;
        sgn = -sgn;
        facn = facn * n;
        FL(12); TR(1);
    }
    return s / 4.0 / PI;
}
static double dsam(double r, double pa, double wvn)
{
    const double eps = F32(1.e-6);
    int n = 1;
    double rm = 100.0, s = 0.0, sgn = -1.0, facn = 1.0;
    while (fabs(rm) > eps) {
        double en = hmo_enxa(n + 1, (wvn * pa / 2.0) * (wvn * pa / 2.0));
        rm = 2.0 * n * sgn * pow(r / pa, 2.0 * n - 1.0) / facn / pa * en;
        s = s + rm;
        n = n + 1;
        sgn = -sgn;
        facn = facn * n;
        FL(15); TR(1);
    }
    return s / 4.0 / PI;
}

/* COEF / DCOEF, FinGreen3D.f90:1168-1340: Linton expansion coefficients and
 * their R-derivatives; cof[0] is the propagating mode, cof[1..m-1] the
 * evanescent ones. */
static void coef(double r, double pa, double a, int m, const double *wvno, double wk, double *cof)
{
    const double eps = F32(1.e-6), step = 100.0;
    double dnt = pa * pa / 4.0, by0, j0, j1, y1;
    FL(2);
    if (fabs(r) < F32(1.e-6)) {
        cof[0] = -hmo_eix(wk * wk * dnt) / 4.0 / PI;
    } else if (r / a <= 0.5) {
        cof[0] = sa0(r, pa, wk);
    } else {
        hmo_jy01b(wk * r, &j0, &j1, &by0, &y1);
        cof[0] = -by0 / 4.0 - aquad(r, wk, 1, 0.0, dnt, eps, 1);
    }
    for (int i = 2; i <= m; i++) {
        if (fabs(r) < F32(1.e-6)) {
            cof[i - 1] = hmo_e1xa(wvno[i - 1] * wvno[i - 1] * dnt) / 4.0 / PI;
        } else if (r / a <= 1.0) {
            cof[i - 1] = sam(r, pa, wvno[i - 1]);
        } else {
            double p1 = dnt + step;
            while (fabs(f1(r, wvno[i - 1], i, p1)) > eps) p1 = p1 + step;
            cof[i - 1] = aquad(r, wvno[i - 1], i, dnt, p1, eps, 1);
        }
    }
}
static void dcoef(double r, double pa, double a, int m, const double *wvno, double wk, double *dcof)
{
    const double eps = F32(1.e-6), step = 100.0;
    double dnt = pa * pa / 4.0, by1, j0, j1, y0;
    FL(2);
    if (fabs(r) < F32(1.e-6)) {
        dcof[0] = 0.0;
    } else if (r / a <= 0.5) {
        dcof[0] = dsa0(r, pa, wk);
    } else {
        hmo_jy01b(wk * r, &j0, &j1, &y0, &by1);
        dcof[0] = wk * by1 / 4.0 - aquad(r, wk, 1, 0.0, dnt, eps, 2);
    }
    for (int i = 2; i <= m; i++) {
        if (fabs(r) < F32(1.e-6)) {
            dcof[i - 1] = 0.0;
        } else if (r / a <= 1.0) {
            dcof[i - 1] = dsam(r, pa, wvno[i - 1]);
        } else {
            double p1 = dnt + step;
            while (fabs(f2(r, wvno[i - 1], i, p1)) > eps) p1 = p1 + step;
            dcof[i - 1] = aquad(r, wvno[i - 1], i, dnt, p1, eps, 2);
        }
    }
}

/* LINTON, FinGreen3D.f90:712-861: Ewald-type splitting (Linton 1999) for
 * R/h < 0.0005, with the parameter A from a cubic fit in V (:741-750). */
static void linton(double r, double zf, double zp, double v, double wk, const double *wvno,
                   double h, double complex g[3])
{
    const double eps = F32(1.e-6);
    const int nt = 20;
    double a, cof[20], dcof[20], ro[4], dror[4], droz[4], errsum[3];
    if (v < 0.02) a = 3324.2 * (v * v * v) - 258.89 * (v * v) + 8.3958 * v + 1.3208;
    else if (v < 0.1) a = -30.303 * (v * v * v) - 2.8138 * (v * v) + 2.7740 * v + 0.05052;
    else if (v < 1.0) a = 1.5699 * (v * v * v) - 2.5075 * (v * v) + 1.8316 * v + 0.1077;
    else a = -0.000069641 * (v * v * v) - 0.0017082 * (v * v) + 0.45566 * v + 0.80551;
    a = fabs(a) / (wk * h);
    double pa = a * h;
    double sr = sqrt(r * r + (zf - zp) * (zf - zp));
    double sr1 = sqrt(r * r + (2.0 * h + zf + zp) * (2.0 * h + zf + zp));
    double sr2 = sqrt(r * r + (zf + zp) * (zf + zp));
    FL(40);
    rchi(r, zf, zp, h, ro, dror, droz);
    double dsrr = r / sr, dsrz = (zf - zp) / sr, dsr1r = r / sr1, dsr1z = (2.0 * h + zf + zp) / sr1;
    double dsr2r = r / sr2, dsr2z = (zf + zp) / sr2;
    FL(12);

    g[0] = g[1] = g[2] = 0.0;
    coef(r, pa, a, nt, wvno, wk, cof);
    dcoef(r, pa, a, nt, wvno, wk, dcof);
    for (int i = 1; i <= nt; i++) {
        double complex km = (i == 1) ? -I * wk : (double complex)wvno[i - 1];
        double complex nm = nmf(h, km);
        double complex c1 = ccos(km * (zf + h)), c2 = ccos(km * (zp + h)), s1 = csin(km * (zf + h));
        double complex rm1 = -cof[i - 1] / nm * c1 * c2;
        double complex rm2 = -dcof[i - 1] / nm * c1 * c2;
        double complex rm3 = cof[i - 1] * km / nm * s1 * c2;
        g[0] += rm1; g[1] += rm2; g[2] += rm3;
        FL(60); TR(6);
        if (i > 2 && cabs(rm1) < eps && cabs(rm2) < eps && cabs(rm3) < eps) break;
    }

    errsum[0] = errsum[1] = errsum[2] = 0.0;
    for (int i = 0; i < 4; i++) {
        double e = erfcc(ro[i] / pa), de = derfcc(ro[i] / pa);
        errsum[0] = errsum[0] + e / (4.0 * PI * ro[i]);
        errsum[1] = errsum[1] + (de / pa * dror[i] * ro[i] - dror[i] * e) / (4.0 * PI * ro[i] * ro[i]);
        errsum[2] = errsum[2] + (de / pa * droz[i] * ro[i] - droz[i] * e) / (4.0 * PI * ro[i] * ro[i]);
        FL(30);
    }
    if (fabs(sr) > F32(1.e-6)) {
        double e = erfcc(sr / pa), de = derfcc(sr / pa), e1 = erfcc(sr1 / pa), de1 = derfcc(sr1 / pa);
        errsum[0] = errsum[0] + (e - 1.0) / (4.0 * PI * sr) + e1 / (4.0 * PI * sr1) - 1.0 / (4.0 * PI * sr2);
        errsum[1] = errsum[1] - (e - 1.0) / (4.0 * PI * (sr * sr)) * dsrr + de / (4.0 * PI * sr * pa) * dsrr
                    - e1 / (4.0 * PI * (sr1 * sr1)) * dsr1r + de1 / (4.0 * PI * sr1 * pa) * dsr1r
                    + dsr2r / (4.0 * PI * (sr2 * sr2));
        errsum[2] = errsum[2] - (e - 1.0) / (4.0 * PI * (sr * sr)) * dsrz + de / (4.0 * PI * sr * pa) * dsrz
                    - e1 / (4.0 * PI * (sr1 * sr1)) * dsr1z + de1 / (4.0 * PI * sr1 * pa) * dsr1z
                    + dsr2z / (4.0 * PI * (sr2 * sr2));
        FL(70);
    } else {
        double e1 = erfcc(sr1 / pa), de1 = derfcc(sr1 / pa);
        errsum[0] = errsum[0] + e1 / (4.0 * PI * sr1) - 1.0 / (4.0 * PI * sr2);
        errsum[1] = errsum[1] - e1 / (4.0 * PI * (sr1 * sr1)) * dsr1r + de1 / (4.0 * PI * sr1 * pa) * dsr1r
                    + dsr2r / (4.0 * PI * (sr2 * sr2));
        errsum[2] = errsum[2] - e1 / (4.0 * PI * (sr1 * sr1)) * dsr1z + de1 / (4.0 * PI * sr1 * pa) * dsr1z
                    + dsr2z / (4.0 * PI * (sr2 * sr2));
        FL(40);
    }

    g[0] = g[0] - errsum[0] - v / 2.0 / PI * aquadf(v, r, zf, zp, h, 0.0, pa / 2.0, eps, 1);
    g[1] = g[1] - errsum[1] + v / 2.0 / PI * aquadf(v, r, zf, zp, h, 0.0, pa / 2.0, eps, 2);
    g[2] = g[2] - errsum[2] - v / 2.0 / PI * aquadf(v, r, zf, zp, h, 0.0, pa / 2.0, eps, 3);

    double bj0, bj1, y0, y1;
    hmo_jy01b(wk * r, &bj0, &bj1, &y0, &y1);
    double complex km = -I * wk, nm = nmf(h, km);
    double chf = cosh(wk * (zf + h)), chp = cosh(wk * (zp + h)), shf = sinh(wk * (zf + h));
    g[0] = g[0] - I / 4.0 / nm * bj0 * chf * chp;
    g[1] = g[1] + I * wk / 4.0 / nm * bj1 * chf * chp;
    g[2] = g[2] - I * wk / 4.0 / nm * bj0 * shf * chp;
    g[0] = -4.0 * PI * g[0];
    g[1] = -4.0 * PI * g[1];
    g[2] = -4.0 * PI * g[2];
    FL(60); TR(3);
}

/* FINGREEN3D_R, FinGreen3D.f90:160-226: region selection on R/h. */
void hmo_fingreen3d_r(double rr, double zf, double zp, double v, const double *wvn, int nk,
                      double h, double complex grn[3], int tag)
{
    double complex fg[3] = {0, 0, 0};
    double wk = wvn[0];
    double rhp = fabs(rr / h);
    double zpq = zf + zp;
    double r = sqrt(rr * rr + (zp - zf) * (zp - zf));
    double r1 = sqrt(rr * rr + zpq * zpq);
    FL(10);
    if (rhp < 0.0005) linton(rr, zf, zp, v, wk, wvn, h, fg);
    else if (rhp < 0.05) pidcock(rr, zf, zp, v, wk, wvn, h, fg);
    else if (rhp < 0.5) eigene(rr, zf, zp, wk, wvn, nk, h, fg);
    else eigen(rr, zf, zp, wk, wvn, h, fg);
    grn[1] = (rr != 0.0) ? fg[1] : 0.0;
    grn[0] = fg[0];
    grn[2] = fg[2];
    if (tag == 2) {
        grn[0] = grn[0] + 1.0 / r1;
        grn[1] = grn[1] - rr / (r1 * r1 * r1);
        grn[2] = grn[2] - zpq / (r1 * r1 * r1);
    } else if (tag == 3) {
        grn[0] = grn[0] + (1.0 / r1 + 1.0 / r);
        grn[1] = grn[1] - (rr / (r1 * r1 * r1) + rr / (r * r * r));
        grn[2] = grn[2] - (zpq / (r1 * r1 * r1) + (zf - zp) / (r * r * r));
    }
}

/* FINGREEN3D, FinGreen3D.f90:81-155. Derivatives are with respect to the
 * source point (xf,yf,zf); V=0 and V=-1 give the zero/infinite-frequency
 * limits. */
void hmo_fingreen3d(double xf, double xp, double yf, double yp, double zf, double zp,
                    double v, const double *wvn, int nk, double h, double complex grn[4], int flag)
{
    double xpq = xf - xp, ypq = yf - yp, zpq = zf + zp;
    double rr = sqrt(xpq * xpq + ypq * ypq);
    double r = sqrt(rr * rr + (zp - zf) * (zp - zf));
    double r1 = sqrt(rr * rr + zpq * zpq);
    FL(14);
    grn[0] = grn[1] = grn[2] = grn[3] = 0.0;
    if (fabs(v + 1.0) < F32(1.e-8) && !(fabs(v) < F32(1.e-8))) {
        if (flag == 1) {
            grn[0] = grn[0] - 1.0 / r1;
            grn[1] = grn[1] + xpq / (r1 * r1 * r1);
            grn[2] = grn[2] + ypq / (r1 * r1 * r1);
            grn[3] = grn[3] + zpq / (r1 * r1 * r1);
        } else {
            grn[0] = grn[0] + (-1.0 / r1 + 1.0 / r);
            grn[1] = grn[1] + (xpq / (r1 * r1 * r1) - xpq / (r * r * r));
            grn[2] = grn[2] + (ypq / (r1 * r1 * r1) - ypq / (r * r * r));
            grn[3] = grn[3] + (zpq / (r1 * r1 * r1) - (zf - zp) / (r * r * r));
        }
        return;
    }
    if (!(fabs(v) < F32(1.e-8))) {
        double complex fg[3];
        hmo_fingreen3d_r(rr, zf, zp, v, wvn, nk, h, fg, flag);
        if (rr > F32(1.e-6)) {
            grn[1] = fg[1] * xpq / rr;
            grn[2] = fg[1] * ypq / rr;
        } else {
            grn[1] = 0.0;
            grn[2] = 0.0;
        }
        grn[0] = fg[0];
        grn[3] = fg[2];
        FL(6);
    }
    if (flag == 1) {
        grn[0] = grn[0] + 1.0 / r1;
        grn[1] = grn[1] - xpq / (r1 * r1 * r1);
        grn[2] = grn[2] - ypq / (r1 * r1 * r1);
        grn[3] = grn[3] - zpq / (r1 * r1 * r1);
        FL(12);
    } else {
        grn[0] = grn[0] + (1.0 / r1 + 1.0 / r);
        grn[1] = grn[1] - (xpq / (r1 * r1 * r1) + xpq / (r * r * r));
        grn[2] = grn[2] - (ypq / (r1 * r1 * r1) + ypq / (r * r * r));
        grn[3] = grn[3] - (zpq / (r1 * r1 * r1) + (zf - zp) / (r * r * r));
        FL(24);
    }
}

/* =============================================================================
 * Infinite-depth Green function -- src/InfGreen_Appr.f90 (Wu, Zhang, Liang,
 * Noblesse 2017 global approximation of the Havelock function). Unlike the
 * reference, the parameters alpha/beta/sigma/rho/d are locals (the reference
 * keeps them as module variables, :89-96, which races under OpenMP).
 * ========================================================================== */
typedef struct { double alpha, beta, sigma, rho, dd; } gf_par;

static double poly(const double *c, int deg, double t)
{   /* Horner in the reference's form: ((c[deg]*t + c[deg-1])*t + ...) + c[0] */
    double p = (c[deg - 1] + c[deg] * t) * t;
    for (int k = deg - 2; k >= 1; k--) p = (c[k] + p) * t;
    FL(2 * deg);
    return c[0] + p;
}
static const double GFA[10] = {+1.21, -13.328, +215.896, -1763.96, +8418.94, -24314.21,
                               +42002.57, -41592.9, 21859.0, -4838.6};       /* :311-315 */
static const double GFB[10] = {+0.938, +5.373, -67.92, +796.534, -4780.77, +17137.74,
                               -36618.81, +44894.06, -29030.24, +7671.22};   /* :340-344 */
static const double GFC[8] = {+1.268, -9.747, +209.653, -1397.89, +5155.67, -9844.35,
                              +9136.4, -3272.62};                            /* :369-372 */
static const double GFD[10] = {+0.632, -40.97, +667.16, -6072.07, +31127.39, -96293.05,
                               +181856.75, -205690.43, +128170.2, -33744.6}; /* :395-400 */
static const double GDA[10] = {+2.948, -24.53, +249.69, -754.85, -1187.71, +16370.75,
                               -48811.41, +68220.87, -46688.0, +12622.25};   /* :508-512 */
static const double GDB[10] = {+1.11, +2.894, -76.765, +1565.35, -11336.19, +44270.15,
                               -97014.11, +118879.26, -76209.82, +19923.28}; /* :537-541 */
static const double GDC[6] = {+14.19, -148.24, +847.8, -2318.58, +3168.35, -1590.27}; /* :566-567 */

/* BesselJ0/J1/Y0/Y1 fits, InfGreen_Appr.f90:583-761 */
static double besj0(double x)
{
    if (x <= 3.0) {
        double y = (x / 3.0) * (x / 3.0);
        FL(14);
        return +0.999999999 + (-2.249999879 + (+1.265623060 + (-0.316394552 + (+0.044460948
               + (-0.003954479 + 0.000212950 * y) * y) * y) * y) * y) * y;
    }
    double y = 3.0 / x, y2 = y * y;
    double f0 = 0.79788454 + (-0.00553897 + (0.00099336 + (-0.00044346 + (0.00020445
                - 0.00004959 * y2) * y2) * y2) * y2) * y2;
    double th = x - 0.25 * PI + (-0.04166592 + (0.00239399 + (-0.00073984 + (0.00031099
                - 0.00007605 * y2) * y2) * y2) * y2) * y;
    FL(30); TR(2);
    return f0 * cos(th) / sqrt(x);
}
static double besj1(double x)
{
    if (x <= 3.0) {
        double y = x / 3.0, y2 = y * y;
        double b = +0.5 + (-0.562499992 + (+0.210937377 + (-0.039550040 + (+0.004447331
                   + (-0.000330547 + 0.000015525 * y2) * y2) * y2) * y2) * y2) * y2;
        FL(15);
        return b * x;
    }
    double y = 3.0 / x, y2 = y * y;
    double f1 = 0.79788459 + (0.01662008 + (-0.00187002 + (0.00068519 + (-0.00029440
                + 0.00006952 * y2) * y2) * y2) * y2) * y2;
    double th = x - 0.75 * PI + (0.12499895 + (-0.00605240 + (0.00135825 + (-0.00049616
                + 0.00011531 * y2) * y2) * y2) * y2) * y;
    FL(30); TR(2);
    return f1 * cos(th) / sqrt(x);
}
static double besy0(double x)
{
    if (x <= 3.0) {
        double q = x / 3.0;
        FL(40); TR(7);
        return (2.0 / PI) * log(x / 2.0) * besj0(x) + 0.367466907 + 0.605593797 * pow(q, 2)
               - 0.743505078 * pow(q, 4) + 0.253005481 * pow(q, 6) - 0.042619616 * pow(q, 8)
               + 0.004285691 * pow(q, 10) - 0.000250716 * pow(q, 12);
    }
    double q = 3.0 / x;
    double f0 = 0.79788454 - 0.00553897 * pow(q, 2) + 0.00099336 * pow(q, 4) - 0.00044346 * pow(q, 6)
                + 0.00020445 * pow(q, 8) - 0.00004959 * pow(q, 10);
    double th = x - 0.25 * PI - 0.04166592 * pow(q, 1) + 0.00239399 * pow(q, 3) - 0.00073984 * pow(q, 5)
                + 0.00031099 * pow(q, 7) - 0.00007605 * pow(q, 9);
    FL(30); TR(13);
    return f0 * sin(th) / sqrt(x);
}
static double besy1(double x)
{
    if (x <= 3.0) {
        double q = x / 3.0;
        FL(40); TR(7);
        return (2.0 / PI) * (log(x / 2.0) * besj1(x) - 1.0 / x) + 0.07373571 * pow(q, 1)
               + 0.72276433 * pow(q, 3) - 0.43885620 * pow(q, 5) + 0.10418264 * pow(q, 7)
               - 0.01340825 * pow(q, 9) + 0.00094249 * pow(q, 11);
    }
    double q = 3.0 / x;
    double f1 = 0.79788459 + 0.01662008 * pow(q, 2) - 0.00187002 * pow(q, 4) + 0.00068519 * pow(q, 6)
                - 0.00029440 * pow(q, 8) + 0.00006952 * pow(q, 10);
    double th = x - 3.0 * PI / 4.0 + 0.12499895 * pow(q, 1) - 0.00605240 * pow(q, 3)
                + 0.00135825 * pow(q, 5) - 0.00049616 * pow(q, 7) + 0.00011531 * pow(q, 9);
    FL(30); TR(13);
    return f1 * sin(th) / sqrt(x);
}
/* StruveH0/H1 fits, InfGreen_Appr.f90:765-872 */
static double struveh0(double x)
{
    if (x <= 3.0) {
        double y = (x / 3.0) * (x / 3.0);
        double s = +1.909859164 + (-1.909855001 + (+0.687514637 + (-0.126164557 + (+0.013828813
                   - 0.000876918 * y) * y) * y) * y) * y;
        FL(14);
        return s * (x / 3.0);
    }
    double y = (3.0 / x) * (3.0 / x);
    double c1 = 2.0 * (0.99999906 + (4.77228920 + (3.85542044 + 0.32303607 * y) * y) * y);
    double c2 = PI * x * (1.0 + (4.88331068 + (4.28957333 + 0.52120508 * y) * y) * y);
    FL(20);
    return c1 / c2 + besy0(x);
}
static double struveh1(double x)
{
    if (x <= 3.0) {
        double y = (x / 3.0) * (x / 3.0);
        FL(14);
        return (+1.909859286 + (-1.145914713 + (+0.294656958 + (-0.042070508 + (+0.003785727
               - 0.000207183 * y) * y) * y) * y) * y) * y;
    }
    double y = (3.0 / x) * (3.0 / x);
    double c1 = 2.0 * (1.00000004 + (3.92205313 + (2.64893033 + 0.27450895 * y) * y) * y);
    double c2 = PI * (1.0 + (3.81095112 + (2.26216956 + 0.10885141 * y) * y) * y);
    FL(20);
    return c1 / c2 + besy1(x);
}

/* HavelockGF, InfGreen_Appr.f90:189-206, with GF_DivParameters :209-224,
 * GF_Func_L0 :227, _Lp :249, _W :277, _Ls :427, _Lsp :450, _Wh :471. */
static void havelock(double hh, double vv, double complex *gf, double complex *gfh)
{
    const double gama = 0.5772156649;
    gf_par q;
    q.dd = sqrt(hh * hh + vv * vv);
    q.alpha = -vv / q.dd;
    q.beta = hh / q.dd;
    q.sigma = hh / (q.dd - vv);
    q.rho = q.dd / (1.0 + q.dd);
    FL(10);
    /* L0 */
    double pp = log(0.5 * (q.dd - vv)) + gama - 2.0 * q.dd * q.dd;
    pp = exp(vv) * pp;
    pp = pp + q.dd * q.dd - vv;
    double a = poly(GFA, 9, q.rho), b = poly(GFB, 9, q.rho), c = poly(GFC, 7, q.rho), d = poly(GFD, 9, q.rho);
    double rr = (1.0 - q.beta) * a;
    rr = rr - q.beta * b;
    rr = rr - q.alpha * c / (1.0 + 6.0 * q.alpha * q.rho * (1.0 - q.rho));
    rr = rr + q.beta * (1.0 - q.beta) * d;
    double lp = q.rho * pow(1.0 - q.rho, 3) * rr;
    double l0 = 2.0 * pp / (1.0 + pow(q.dd, 3)) + 2.0 * lp;
    FL(40); TR(4);
    /* W */
    double h0 = struveh0(hh), j0 = besj0(hh);
    double complex w = 2.0 * PI * (h0 - I * j0) * exp(vv);
    /* Ls */
    double ps = (q.beta + hh) / (q.dd - vv);
    ps = ps - 2.0 * q.beta + 2.0 * exp(vv) * q.dd - hh;
    double qs = exp(-q.dd) * (1.0 - q.beta);
    qs = qs * (1.0 + q.dd / (1.0 + pow(q.dd, 3)));
    double da = poly(GDA, 9, q.rho), db = poly(GDB, 9, q.rho), dc = poly(GDC, 5, q.rho);
    double r2 = q.beta * da;
    r2 = r2 - (1.0 - q.alpha) * db;
    r2 = r2 + q.beta * (1.0 - q.beta) * q.rho * (1.0 - 2.0 * q.rho) * dc;
    double lsp = q.rho * pow(1.0 - q.rho, 3) * r2;
    double ls = 2.0 * ps / (1.0 + pow(q.dd, 3)) - 4.0 * qs + 2.0 * lsp;
    /* Wh */
    double h1 = struveh1(hh), j1 = besj1(hh);
    double complex wh = 2.0 * PI * (2.0 / PI - h1 + I * j1) * exp(vv);
    FL(60); TR(7);
    *gf = l0 + w;
    *gfh = ls + wh;
}

/* INFGREEN3D, InfGreen_Appr.f90:104-186. */
void hmo_infgreen3d(double xf, double xp, double yf, double yp, double zf, double zp,
                    double v, double complex grn[4], int flag)
{
    double xpq = xf - xp, ypq = yf - yp, zpq = zp + zf;
    double rr = sqrt(xpq * xpq + ypq * ypq);
    double r = sqrt(rr * rr + (zp - zf) * (zp - zf));
    double r1 = sqrt(rr * rr + zpq * zpq);
    double complex gf, gfh;
    FL(14);
    grn[0] = grn[1] = grn[2] = grn[3] = 0.0;
    havelock(v * rr, v * zpq, &gf, &gfh);
    FL(2);
    if (fabs(v) < F32(1.e-8)) {
        if (flag == 1) {
            grn[0] += 1.0 / r1;
            grn[1] += -xpq / (r1 * r1 * r1);
            grn[2] += -ypq / (r1 * r1 * r1);
            grn[3] += -zpq / (r1 * r1 * r1);
        } else {
            grn[0] = grn[0] + 1.0 / r1 + 1.0 / r;
            grn[1] = grn[1] - xpq / (r1 * r1 * r1) - xpq / (r * r * r);
            grn[2] = grn[2] - ypq / (r1 * r1 * r1) - ypq / (r * r * r);
            grn[3] = grn[3] - zpq / (r1 * r1 * r1) - (zf - zp) / (r * r * r);
        }
    } else if (fabs(v + 1.0) < F32(1.e-8)) {
        if (flag == 1) {
            grn[0] = grn[0] - 1.0 / r1;
            grn[1] = grn[1] + xpq / (r1 * r1 * r1);
            grn[2] = grn[2] + ypq / (r1 * r1 * r1);
            grn[3] = grn[3] + zpq / (r1 * r1 * r1);
        } else {
            grn[0] = grn[0] - 1.0 / r1 + 1.0 / r;
            grn[1] = grn[1] + xpq / (r1 * r1 * r1) - xpq / (r * r * r);
            grn[2] = grn[2] + ypq / (r1 * r1 * r1) - ypq / (r * r * r);
            grn[3] = grn[3] + zpq / (r1 * r1 * r1) - (zf - zp) / (r * r * r);
        }
    } else {
        if (rr > F32(1.e-6)) {
            grn[1] = -(v * v) * gfh * xpq / rr;
            grn[2] = -(v * v) * gfh * ypq / rr;
        } else {
            grn[1] = 0.0;
            grn[2] = 0.0;
        }
        grn[0] = -v * gf;
        grn[3] = -(v * v) * gf + 2.0 * v / r1;
        if (flag == 1) {
            grn[0] = grn[0] + 1.0 / r1;
            grn[1] = grn[1] - xpq / (r1 * r1 * r1);
            grn[2] = grn[2] - ypq / (r1 * r1 * r1);
            grn[3] = grn[3] - zpq / (r1 * r1 * r1);
        } else {
            grn[0] = grn[0] + 1.0 / r1 + 1.0 / r;
            grn[1] = grn[1] - xpq / (r1 * r1 * r1) - xpq / (r * r * r);
            grn[2] = grn[2] - ypq / (r1 * r1 * r1) - ypq / (r * r * r);
            grn[3] = grn[3] - zpq / (r1 * r1 * r1) - (zf - zp) / (r * r * r);
        }
        FL(40);
    }
}

/* =============================================================================
 * Rankine panel integrals -- src/SingularIntgr.f90 (Newman 1986)
 * ========================================================================== */

/* SGLINTBD_TRI, SingularIntgr.f90:44-244: a quadrilateral given to the TRI
 * routine is split into (1,2,3) and (1,3,4) (:221-226). */
static void sglint_tri(int ncn, const double *xq, const double xf[3], double *sij, double dij[3])
{
    double xv[3][3];
    *sij = 0.0;
    dij[0] = dij[1] = dij[2] = 0.0;
    for (int k = 0; k < 3; k++) { xv[0][k] = xq[k]; xv[1][k] = xq[3 + k]; xv[2][k] = xq[6 + k]; }
    for (int itno = 1; itno <= 2; itno++) {
        double p12[3], p13[3], po[3], po3[3], ui[3], uj[3], uk[3], pof[3], po1[3], po2[3];
        double ksi[4], eta[4];
        for (int k = 0; k < 3; k++) { p12[k] = xv[1][k] - xv[0][k]; p13[k] = xv[2][k] - xv[0][k]; }
        double alfa = (p12[0] * p13[0] + p12[1] * p13[1] + p12[2] * p13[2])
                      / (p12[0] * p12[0] + p12[1] * p12[1] + p12[2] * p12[2]);
        for (int k = 0; k < 3; k++) po[k] = xv[0][k] + alfa * p12[k];
        for (int k = 0; k < 3; k++) po3[k] = xv[2][k] - po[k];
        double l12 = sqrt(p12[0] * p12[0] + p12[1] * p12[1] + p12[2] * p12[2]);
        double l3 = sqrt(po3[0] * po3[0] + po3[1] * po3[1] + po3[2] * po3[2]);
        for (int k = 0; k < 3; k++) { ui[k] = p12[k] / l12; uj[k] = po3[k] / l3; }
        uk[0] = ui[1] * uj[2] - ui[2] * uj[1];
        uk[1] = ui[2] * uj[0] - ui[0] * uj[2];
        uk[2] = ui[0] * uj[1] - ui[1] * uj[0];
        for (int k = 0; k < 3; k++) pof[k] = xf[k] - po[k];
        double x = ui[0] * pof[0] + ui[1] * pof[1] + ui[2] * pof[2];
        double y = uj[0] * pof[0] + uj[1] * pof[1] + uj[2] * pof[2];
        double z = uk[0] * pof[0] + uk[1] * pof[1] + uk[2] * pof[2];
        for (int k = 0; k < 3; k++) { po1[k] = xv[0][k] - po[k]; po2[k] = xv[1][k] - po[k]; }
        ksi[0] = -copysign(1.0, alfa) * sqrt(po1[0] * po1[0] + po1[1] * po1[1] + po1[2] * po1[2]);
        eta[0] = 0.0;
        ksi[1] = 0.0;
        eta[1] = l3;
        ksi[2] = sqrt(po2[0] * po2[0] + po2[1] * po2[1] + po2[2] * po2[2]);
        eta[2] = 0.0;
        ksi[3] = ksi[0];
        eta[3] = eta[0];
        double st = 0.0, sxt = 0.0, syt = 0.0, szt = 0.0;
        FL(70); TR(0);
        for (int i = 0; i < 3; i++) {
            FL(40); TR(4);
            double dksi = ksi[i + 1] - ksi[i], deta = eta[i + 1] - eta[i];
            double r1 = sqrt((ksi[i] - x) * (ksi[i] - x) + (eta[i] - y) * (eta[i] - y) + z * z);
            double r2 = sqrt((ksi[i + 1] - x) * (ksi[i + 1] - x) + (eta[i + 1] - y) * (eta[i + 1] - y) + z * z);
            double l = sqrt(dksi * dksi + deta * deta);
            double b1 = deta * ((ksi[i] - x) * (ksi[i] - x) + z * z) - dksi * (ksi[i] - x) * (eta[i] - y);
            double a1 = r1 * z * dksi;
            double b2 = deta * ((ksi[i + 1] - x) * (ksi[i + 1] - x) + z * z) - dksi * (ksi[i + 1] - x) * (eta[i + 1] - y);
            double a2 = r2 * z * dksi;
            if (!(fabs(z) < 1.0e-6)) szt = szt + atan2((b1 * a2 - b2 * a1), (a1 * a2 + b1 * b2));
            if (fabs(r1 + r2 - l) < 1.0e-6) continue;
            double lgrn = log((r1 + r2 + l) / (r1 + r2 - l));
            sxt = sxt - deta / l * lgrn;
            syt = syt + dksi / l * lgrn;
            st = st + (deta * (x - ksi[i]) - dksi * (y - eta[i])) / l * lgrn;
        }
        st = st - z * szt;
        *sij = *sij + st;
        for (int k = 0; k < 3; k++) dij[k] = dij[k] + ui[k] * sxt + uj[k] * syt + uk[k] * szt;
        if (ncn == 3 || itno == 2) break;
        for (int k = 0; k < 3; k++) { xv[0][k] = xq[k]; xv[1][k] = xq[6 + k]; xv[2][k] = xq[9 + k]; }
    }
}

/* SGLINTBD_QUAD, SingularIntgr.f90:252-445: flat-quadrilateral form on the
 * panel's mean plane; edges shorter than 1e-8 are skipped (:416). */
static void sglint_quad(const double *xq, const double xf[3], double *sij, double dij[3])
{
    double xv[4][3], po[3], pov[4][3], pon[4][3], un[3], ui[3], uj[3], uk[3], pof[3];
    double ksi[5], eta[5];
    for (int v = 0; v < 4; v++) for (int k = 0; k < 3; k++) xv[v][k] = xq[3 * v + k];
    for (int k = 0; k < 3; k++) po[k] = (xv[0][k] + xv[1][k] + xv[2][k] + xv[3][k]) / 4.0;
    for (int k = 0; k < 3; k++) { pov[0][k] = xv[0][k] - po[k]; pov[1][k] = xv[1][k] - po[k]; }
    un[0] = pov[0][1] * pov[1][2] - pov[0][2] * pov[1][1];
    un[1] = pov[0][2] * pov[1][0] - pov[0][0] * pov[1][2];
    un[2] = pov[0][0] * pov[1][1] - pov[0][1] * pov[1][0];
    double lp = sqrt(pov[0][0] * pov[0][0] + pov[0][1] * pov[0][1] + pov[0][2] * pov[0][2]);
    double ln = sqrt(un[0] * un[0] + un[1] * un[1] + un[2] * un[2]);
    for (int k = 0; k < 3; k++) { ui[k] = pov[0][k] / lp; uk[k] = un[k] / ln; }
    uj[0] = uk[1] * ui[2] - uk[2] * ui[1];
    uj[1] = uk[2] * ui[0] - uk[0] * ui[2];
    uj[2] = uk[0] * ui[1] - uk[1] * ui[0];
    for (int k = 0; k < 3; k++) pof[k] = xf[k] - po[k];
    double x = ui[0] * pof[0] + ui[1] * pof[1] + ui[2] * pof[2];
    double y = uj[0] * pof[0] + uj[1] * pof[1] + uj[2] * pof[2];
    double z = uk[0] * pof[0] + uk[1] * pof[1] + uk[2] * pof[2];
    for (int k = 0; k < 3; k++) { pov[2][k] = xv[2][k] - po[k]; pov[3][k] = xv[3][k] - po[k]; }
    for (int i = 0; i < 4; i++) {
        double dk = uk[0] * pov[i][0] + uk[1] * pov[i][1] + uk[2] * pov[i][2];
        for (int j = 0; j < 3; j++) pon[i][j] = pov[i][j] - dk * uk[j];
    }
    for (int i = 0; i < 4; i++) {
        ksi[i] = ui[0] * pon[i][0] + ui[1] * pon[i][1] + ui[2] * pon[i][2];
        eta[i] = uj[0] * pon[i][0] + uj[1] * pon[i][1] + uj[2] * pon[i][2];
    }
    ksi[4] = ksi[0];
    eta[4] = eta[0];
    double st = 0.0, sxt = 0.0, syt = 0.0, szt = 0.0;
    FL(110);
    for (int i = 0; i < 4; i++) {
        FL(40); TR(4);
        double dksi = ksi[i + 1] - ksi[i], deta = eta[i + 1] - eta[i];
        double r1 = sqrt((ksi[i] - x) * (ksi[i] - x) + (eta[i] - y) * (eta[i] - y) + z * z);
        double r2 = sqrt((ksi[i + 1] - x) * (ksi[i + 1] - x) + (eta[i + 1] - y) * (eta[i + 1] - y) + z * z);
        double l = sqrt(dksi * dksi + deta * deta);
        if (fabs(l) < 1.0e-8) continue;
        double b1 = deta * ((ksi[i] - x) * (ksi[i] - x) + z * z) - dksi * (ksi[i] - x) * (eta[i] - y);
        double a1 = r1 * z * dksi;
        double b2 = deta * ((ksi[i + 1] - x) * (ksi[i + 1] - x) + z * z) - dksi * (ksi[i + 1] - x) * (eta[i + 1] - y);
        double a2 = r2 * z * dksi;
        if (!(fabs(z) < 1.0e-6)) szt = szt + atan2((b1 * a2 - b2 * a1), (a1 * a2 + b1 * b2));
        if (fabs(r1 + r2 - l) < 1.0e-6) continue;
        double lgrn = log((r1 + r2 + l) / (r1 + r2 - l));
        sxt = sxt - deta / l * lgrn;
        syt = syt + dksi / l * lgrn;
        st = st - (deta * (x - ksi[i]) - dksi * (y - eta[i])) / l * lgrn;
    }
    st = st + z * szt;
    *sij = st;
    for (int k = 0; k < 3; k++) dij[k] = -(ui[k] * sxt + uj[k] * syt + uk[k] * szt);
}

void hmo_sglintbd(int ncn, const double verts[12], const double xf[3], double *sij, double dij[3])
{
    if (ncn == 3) sglint_tri(3, verts, xf, sij, dij);
    else sglint_quad(verts, xf, sij, dij);
}

/* =============================================================================
 * Panel geometry -- src/NormalProcess.f90 (CalPanelCentre :134, CalPanelArea
 * :171 via CalDeltaArea_Improved :373, CalPanelChartLength :401,
 * CalTransNormals :74, CalRotNormals :46)
 * ========================================================================== */
static double tri_area(const double *p1, const double *p2, const double *p3)
{
    double v21[3], v23[3], v13[3];
    for (int k = 0; k < 3; k++) { v21[k] = p1[k] - p2[k]; v23[k] = p3[k] - p2[k]; v13[k] = p3[k] - p1[k]; }
    double a = sqrt(v21[0] * v21[0] + v21[1] * v21[1] + v21[2] * v21[2]);
    double b = sqrt(v13[0] * v13[0] + v13[1] * v13[1] + v13[2] * v13[2]);
    double c = sqrt(v23[0] * v23[0] + v23[1] * v23[1] + v23[2] * v23[2]);
    double l1 = a + (b + c), l2 = c - (a - b), l3 = c + (a - b), l4 = a + (b - c);
    return sqrt(l1 * l2 * l3 * l4) / 4.0;
}

void hmo_panel_geometry(int nelem, const int *ncn, const int *ncon, const double *xyz,
                        const double rg[3], double *xyz_p, double *ds, double *pnsz, double *dxyz_p)
{
    for (int e = 0; e < nelem; e++) {
        const double *p[4];
        int nc = ncn[e];
        for (int v = 0; v < nc; v++) p[v] = xyz + 3 * (ncon[4 * e + v] - 1);
        double c[3] = {0, 0, 0};
        for (int v = 0; v < nc; v++) for (int k = 0; k < 3; k++) c[k] += p[v][k];
        for (int k = 0; k < 3; k++) xyz_p[3 * e + k] = c[k] / nc;
        double a = tri_area(p[0], p[1], p[2]);
        if (nc == 4) a = a + tri_area(p[0], p[3], p[2]);
        ds[e] = a;
        double m = 0.0;
        for (int v = 0; v < nc; v++) {
            const double *q0 = p[v], *q1 = p[(v + 1) % nc];
            double d[3] = {q1[0] - q0[0], q1[1] - q0[1], q1[2] - q0[2]};
            double l = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (l > m) m = l;
        }
        pnsz[e] = m;
        double un[3];
        if (nc == 3) {
            double v21[3], v23[3];
            for (int k = 0; k < 3; k++) { v21[k] = p[0][k] - p[1][k]; v23[k] = p[2][k] - p[1][k]; }
            un[0] = v21[1] * v23[2] - v21[2] * v23[1];
            un[1] = v21[2] * v23[0] - v21[0] * v23[2];
            un[2] = v21[0] * v23[1] - v21[1] * v23[0];
        } else {
            double v13[3], v24[3];
            for (int k = 0; k < 3; k++) { v13[k] = p[2][k] - p[0][k]; v24[k] = p[3][k] - p[1][k]; }
            un[0] = v24[1] * v13[2] - v24[2] * v13[1];
            un[1] = v24[2] * v13[0] - v24[0] * v13[2];
            un[2] = v24[0] * v13[1] - v24[1] * v13[0];
        }
        double aun = sqrt(un[0] * un[0] + un[1] * un[1] + un[2] * un[2]);
        double *n = dxyz_p + 6 * e;
        n[0] = un[0] / aun; n[1] = un[1] / aun; n[2] = un[2] / aun;
        const double *x = xyz_p + 3 * e;
        n[3] = (x[1] - rg[1]) * n[2] - (x[2] - rg[2]) * n[1];
        n[4] = (x[2] - rg[2]) * n[0] - (x[0] - rg[0]) * n[2];
        n[5] = (x[0] - rg[0]) * n[1] - (x[1] - rg[1]) * n[0];
    }
}

/* =============================================================================
 * Incident wave -- src/PatclVelct.f90 (VINP :42-76, DINP :83-123). VINP treats
 * H>500 as deep water, DINP does not (kept as in the reference).
 * ========================================================================== */
typedef struct { double v, w1, wk, h, amp, g; } wave_t;

static double complex vinp(const wave_t *w, double x, double y, double z, double xw, double yw, double sita)
{
    double w0, wk0;
    if (fabs(w->v) < F32(1.e-8)) { w0 = F32(1.e-20); wk0 = F32(1.e-20); }
    else if (fabs(w->v + 1.0) < F32(1.e-8)) { w0 = F32(1.e+20); wk0 = F32(1.e+20); }
    else { w0 = w->w1; wk0 = w->wk; }
    double wkx = wk0 * ((x - xw) * cos(sita) + (y - yw) * sin(sita));
    double complex r;
    if (w->h < 0.0 || w->h > 500.0) {
        r = -I * w->amp * w->g / w0 * exp(wk0 * z) * cexp(I * wkx);
    } else {
        double dum = -w->amp * w0 / (wk0 * sinh(wk0 * w->h));
        r = I * dum * cosh(wk0 * (z + w->h)) * cexp(I * wkx);
    }
    return r / I;
}

static void dinp(const wave_t *w, double x, double y, double z, double xw, double yw, double sita,
                 double complex *dx, double complex *dy, double complex *dz)
{
    double w0, wk0;
    if (fabs(w->v) < F32(1.e-8)) { w0 = F32(1.e-20); wk0 = F32(1.e-20); }
    else if (fabs(w->v + 1.0) < F32(1.e-8)) { w0 = F32(1.e+20); wk0 = F32(1.e+20); }
    else { w0 = w->w1; wk0 = w->wk; }
    double wkx = wk0 * ((x - xw) * cos(sita) + (y - yw) * sin(sita));
    double complex e = cexp(I * wkx);
    if (w->h < 0.0) {
        *dx = w0 * w->amp * exp(wk0 * z) * e * cos(sita);
        *dy = w0 * w->amp * exp(wk0 * z) * e * sin(sita);
        *dz = -I * w0 * w->amp * exp(wk0 * z) * e;
    } else {
        double thkdh = cosh(wk0 * (z + w->h)) / sinh(wk0 * w->h);
        double thkdv = sinh(wk0 * (z + w->h)) / sinh(wk0 * w->h);
        *dx = w0 * w->amp * thkdh * e * cos(sita);
        *dy = w0 * w->amp * thkdh * e * sin(sita);
        *dz = -I * w0 * w->amp * thkdv * e;
    }
    *dx = *dx / I;
    *dy = *dy / I;
    *dz = *dz / I;
}

/* =============================================================================
 * Dense complex LU -- the textbook right-looking partial-pivot algorithm of
 * LAPACK ZGETRF/ZGETRS (pivot = first max of |re|+|im|, as izamax).
 * ========================================================================== */
int hmo_zgetrf(int n, double complex *a, int lda, int *ipiv)
{
    int info = 0;
    for (int j = 0; j < n; j++) {
        int p = j;
        double best = -1.0;
        for (int i = j; i < n; i++) {
            double m = fabs(creal(a[i + (size_t)j * lda])) + fabs(cimag(a[i + (size_t)j * lda]));
            if (m > best) { best = m; p = i; }
        }
        ipiv[j] = p + 1;
        if (best == 0.0) { if (!info) info = j + 1; continue; }
        if (p != j)
            for (int k = 0; k < n; k++) {
                double complex t = a[j + (size_t)k * lda];
                a[j + (size_t)k * lda] = a[p + (size_t)k * lda];
                a[p + (size_t)k * lda] = t;
            }
        double complex rp = 1.0 / a[j + (size_t)j * lda];
        for (int i = j + 1; i < n; i++) a[i + (size_t)j * lda] *= rp;
        for (int k = j + 1; k < n; k++) {
            double complex u = a[j + (size_t)k * lda];
            if (u == 0.0) continue;
            double complex *ck = a + (size_t)k * lda, *cj = a + (size_t)j * lda;
            for (int i = j + 1; i < n; i++) ck[i] -= cj[i] * u;
        }
    }
    return info;
}

void hmo_zgetrs(int n, int nrhs, const double complex *a, int lda, const int *ipiv,
                double complex *b, int ldb)
{
    for (int c = 0; c < nrhs; c++) {
        double complex *x = b + (size_t)c * ldb;
        for (int i = 0; i < n; i++) {
            int p = ipiv[i] - 1;
            if (p != i) { double complex t = x[i]; x[i] = x[p]; x[p] = t; }
        }
        for (int k = 0; k < n; k++) {
            double complex xk = x[k];
            if (xk == 0.0) continue;
            const double complex *col = a + (size_t)k * lda;
            for (int i = k + 1; i < n; i++) x[i] -= col[i] * xk;
        }
        for (int k = n - 1; k >= 0; k--) {
            const double complex *col = a + (size_t)k * lda;
            x[k] = x[k] / col[k];
            double complex xk = x[k];
            for (int i = 0; i < k; i++) x[i] -= col[i] * xk;
        }
    }
}

/* =============================================================================
 * One panel-frequency system.
 *   CalWaveProperts INFT=3   ImplementSubs.f90:96-105
 *   CALGREEN_MULTI(_IRR)     CalGreenFuncMulti.f90:112-318
 *   BODINT/RBC/DBC           BodyIntgrMulti.f90:48-173, BodyIntgr_irr.f90:49-236
 *   ASSB_LEFT/RBC/DBC        AssbMatxMulti.f90:53-251, AssbMatx_irr.f90:53-466
 *   RADIATION/DIFFRACTION    AssbMatxMulti.f90:257-383
 *   RFORCE/EFORCE_MULTI      PotentWavForceMulti.f90:45-226
 * NSYS=1 (no geometric symmetry), ISOL=1 (total potential in the excitation).
 * ========================================================================== */
static int solve_wave(const hmo_problem *p, wave_t w, const double *wvn, hmo_result *out);

int hmo_solve_frequency(const hmo_problem *p, double w1, hmo_result *out)
{
    wave_t w;
    double wvn[200];
    w.w1 = w1;
    w.v = w1 * w1 / p->g;
    w.h = p->h;
    w.amp = p->amp;
    w.g = p->g;
    if (p->h > 0.0) {
        hmo_dispersion(wvn, 200, w1, p->h, p->g_disp);
        w.wk = wvn[0];
    } else {
        w.wk = w.v;
        for (int k = 0; k < 200; k++) wvn[k] = 0.0;
    }
    return solve_wave(p, w, wvn, out);
}

/* Zero (which = 0) and infinite (which = 1) wave frequency, CalWaveProperts
 * with SYBO = 1 (ImplementSubs.f90:45-70): V = WK = W1 = 0 or -1; DISPERSION is
 * not called (the limits of FINGREEN3D and INFGREEN3D do not read WVN). */
int hmo_solve_limit(const hmo_problem *p, int which, hmo_result *out)
{
    wave_t w;
    double wvn[200];
    const double val = which ? -1.0 : 0.0;
    w.w1 = w.v = w.wk = val;
    w.h = p->h;
    w.amp = p->amp;
    w.g = p->g;
    for (int k = 0; k < 200; k++) wvn[k] = 0.0;
    return solve_wave(p, w, wvn, out);
}

static int solve_wave(const hmo_problem *p, wave_t w, const double *wvn, hmo_result *out)
{
    const int n = p->n, ni = p->ni, nt = n + ni, nb = p->nbody, m = 6 * nb, nbeta = p->nbeta;
    const int ncol = m + nbeta;

    double complex *A = calloc((size_t)nt * n, sizeof(double complex));
    double complex *B = calloc((size_t)nt * ncol, sizeof(double complex));
    double complex *q = calloc((size_t)n * nbeta, sizeof(double complex));
    int *body_of = malloc(sizeof(int) * n);
    for (int b = 0; b < nb; b++)
        for (int j = p->body_start[b]; j < p->body_start[b + 1]; j++) body_of[j] = b;

    /* incident normal velocity of every source panel, per heading (DINP·n) */
    for (int b = 0; b < nbeta; b++)
        for (int j = 0; j < n; j++) {
            double complex dx, dy, dz;
            const double *x = p->xyz_p + 3 * j, *nv = p->dxyz_p + 6 * j;
            dinp(&w, x[0], x[1], x[2], p->xw[0], p->xw[1], p->beta[b], &dx, &dy, &dz);
            q[j + (size_t)b * n] = 1.0 * dx * nv[0] + 1.0 * dy * nv[1] + dz * nv[2];
        }

    for (int i = 0; i < nt; i++) {
        const double *xp = (i < n) ? p->xyz_p + 3 * i : p->ixyz_p + 3 * (i - n);
        for (int j = 0; j < n; j++) {
            const double *xq = p->xyz_p + 3 * j, *env = p->dxyz_p + 6 * j;
            double d0 = xp[0] - xq[0], d1 = xp[1] - xq[1], d2 = xp[2] - xq[2];
            double dist = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            int flag = dist <= 50.0 * p->pnsz[j];
            double sij = 0.0, dij[3] = {0, 0, 0};
            if (flag) {
                double verts[12] = {0};
                for (int v = 0; v < p->ncn[j]; v++)
                    for (int k = 0; k < 3; k++) verts[3 * v + k] = p->xyz[3 * (p->ncon[4 * j + v] - 1) + k];
                hmo_sglintbd(p->ncn[j], verts, xp, &sij, dij);
            }
            double complex grn[4];
            if (p->h < 0.0) hmo_infgreen3d(xq[0], xp[0], xq[1], xp[1], xq[2], xp[2], w.v, grn, flag);
            else hmo_fingreen3d(xq[0], xp[0], xq[1], xp[1], xq[2], xp[2], w.v, wvn, 200, p->h, grn, flag);
            double ear = p->ds[j];
            double complex tindp, dum;
            if (flag) {
                tindp = (dij[0] + grn[1] * ear) * env[0] + (dij[1] + grn[2] * ear) * env[1]
                        + (dij[2] + grn[3] * ear) * env[2];
                dum = sij + grn[0] * ear;
            } else {
                tindp = (grn[1] * env[0] + grn[2] * env[1] + grn[3] * env[2]) * ear;
                dum = grn[0] * ear;
            }
            A[i + (size_t)j * nt] = ((i == j) ? 2.0 * PI : 0.0) + tindp;
            if (out->smat) out->smat[i + (size_t)j * nt] = dum;
            int b = body_of[j];
            for (int d = 0; d < 6; d++) B[i + (size_t)(6 * b + d) * nt] += dum * env[d];
            for (int h = 0; h < nbeta; h++) B[i + (size_t)(m + h) * nt] += -(dum * q[j + (size_t)h * n]);
        }
    }
    if (out->amat) memcpy(out->amat, A, sizeof(double complex) * (size_t)nt * n);
    if (out->rhs) memcpy(out->rhs, B, sizeof(double complex) * (size_t)nt * ncol);

    /* normal equations for the irregular-frequency removal (AssbMatx_irr.f90:151-169,
     * :300-318, :449-463): C = A^T A, D = A^T B (plain transpose) */
    double complex *C, *X;
    if (ni > 0) {
        C = calloc((size_t)n * n, sizeof(double complex));
        X = calloc((size_t)n * ncol, sizeof(double complex));
        for (int jj = 0; jj < n; jj++)
            for (int ii = 0; ii < n; ii++) {
                double complex s = 0.0;
                for (int k = 0; k < nt; k++) s += A[k + (size_t)ii * nt] * A[k + (size_t)jj * nt];
                C[ii + (size_t)jj * n] = s;
            }
        for (int c = 0; c < ncol; c++)
            for (int ii = 0; ii < n; ii++) {
                double complex s = 0.0;
                for (int k = 0; k < nt; k++) s += B[k + (size_t)c * nt] * A[k + (size_t)ii * nt];
                X[ii + (size_t)c * n] = s;
            }
    } else {
        C = A;
        X = malloc(sizeof(double complex) * (size_t)n * ncol);
        for (int c = 0; c < ncol; c++) memcpy(X + (size_t)c * n, B + (size_t)c * nt, sizeof(double complex) * n);
    }
    int *ipiv = malloc(sizeof(int) * n);
    int info = hmo_zgetrf(n, C, n, ipiv);
    hmo_zgetrs(n, ncol, C, n, ipiv, X, n);
    if (out->mxpot) memcpy(out->mxpot, X, sizeof(double complex) * (size_t)n * ncol);

    /* force reductions */
    for (int c = 0; c < m; c++)
        for (int r = 0; r < m; r++) { out->amas[r + (size_t)c * m] = 0.0; out->bdmp[r + (size_t)c * m] = 0.0; }
    for (int bm = 0; bm < nb; bm++) {
        for (int c = 0; c < m; c++)
            for (int i = p->body_start[bm]; i < p->body_start[bm + 1]; i++) {
                double complex phi = X[i + (size_t)c * n];
                double complex rphi = I * p->rho * phi * p->ds[i];
                double complex iphi = I * p->rho * w.w1 * phi * p->ds[i];
                for (int d = 0; d < 6; d++) {
                    out->amas[(6 * bm + d) + (size_t)c * m] += cimag(rphi * 1.0 * p->dxyz_p[6 * i + d]);
                    out->bdmp[(6 * bm + d) + (size_t)c * m] -= creal(iphi * 1.0 * p->dxyz_p[6 * i + d]);
                }
            }
        for (int h = 0; h < nbeta; h++) {
            double complex force[6] = {0, 0, 0, 0, 0, 0};
            for (int i = p->body_start[bm]; i < p->body_start[bm + 1]; i++) {
                const double *x = p->xyz_p + 3 * i;
                double complex phi = (X[i + (size_t)(m + h) * n]
                                      + vinp(&w, x[0], x[1], x[2], p->xw[0], p->xw[1], p->beta[h])) * p->ds[i];
                for (int d = 0; d < 6; d++) force[d] += phi * 1.0 * p->dxyz_p[6 * i + d];
            }
            for (int d = 0; d < 6; d++) out->exfc[(6 * bm + d) + (size_t)h * m] = I * w.w1 * p->rho * force[d];
        }
    }
    if (ni > 0) { free(C); }
    free(X); free(A); free(B); free(q); free(body_of); free(ipiv);
    return info;
}

/* =============================================================================
 * Field-point potential, pressure and elevation -- src/PressureElevationMulti.f90
 * CalPotentialMulti :55-272, CalPressureMulti :305-324, CalElevationMulti
 * :329-348, for every radiation mode (column 6 b + d: body b, dof d; the
 * reference's PHIDOF = b + 1, MD = d + 1) and heading, NSYS = 1, ISOL = 1, after
 * the frequency solve. pot, prs, elv: nfp x (6 nbody + nbeta), column-major.
 * ========================================================================== */
int hmo_field_points(const hmo_problem *p, double w1, int nfp, const double *xfp, double complex *pot,
                     double complex *prs, double complex *elv)
{
    const int n = p->n, nb = p->nbody, m = 6 * nb, nbeta = p->nbeta, ncol = m + nbeta;
    wave_t w;
    double wvn[200];
    w.w1 = w1;
    w.v = w1 * w1 / p->g;
    w.h = p->h;
    w.amp = p->amp;
    w.g = p->g;
    if (p->h > 0.0) {
        hmo_dispersion(wvn, 200, w1, p->h, p->g_disp);
        w.wk = wvn[0];
    } else {
        w.wk = w.v;
        for (int k = 0; k < 200; k++) wvn[k] = 0.0;
    }
    double *amas = calloc((size_t)m * m, sizeof(double)), *bdmp = calloc((size_t)m * m, sizeof(double));
    double complex *exfc = calloc((size_t)m * nbeta, sizeof(double complex));
    double complex *mx = calloc((size_t)n * ncol, sizeof(double complex));
    hmo_result r = {amas, bdmp, exfc, NULL, NULL, NULL, mx};
    int info = solve_wave(p, w, wvn, &r);
    const int limit = fabs(w.v) < F32(1.e-8) || fabs(w.v + 1.0) < F32(1.e-8);
    for (int ip = 0; ip < nfp; ip++) {
        const double *xp = xfp + 3 * ip;
        double complex *acc = calloc((size_t)ncol, sizeof(double complex));
        for (int j = 0; j < n; j++) {
            const double *xq = p->xyz_p + 3 * j, *env = p->dxyz_p + 6 * j;
            double d0 = xp[0] - xq[0], d1 = xp[1] - xq[1], d2 = xp[2] - xq[2];
            int flag = sqrt(d0 * d0 + d1 * d1 + d2 * d2) <= 50.0 * p->pnsz[j];
            double rk[4] = {0, 0, 0, 0};
            if (flag) {                 /* SGLINTBD_MULTI_*2 at the field point, IRR = 1 */
                double verts[12] = {0};
                for (int v = 0; v < p->ncn[j]; v++)
                    for (int k = 0; k < 3; k++) verts[3 * v + k] = p->xyz[3 * (p->ncon[4 * j + v] - 1) + k];
                hmo_sglintbd(p->ncn[j], verts, xp, &rk[0], rk + 1);
            }
            double complex grn[4];
            if (p->h < 0.0) hmo_infgreen3d(xq[0], xp[0], xq[1], xp[1], xq[2], xp[2], w.v, grn, flag);
            else hmo_fingreen3d(xq[0], xp[0], xq[1], xp[1], xq[2], xp[2], w.v, wvn, 200, p->h, grn, flag);
            const double ear = p->ds[j];
            double complex dum1, dum2;
            if (flag) {
                dum1 = rk[0] + grn[0] * ear;
                dum2 = (rk[1] + grn[1] * ear) * env[0] + (rk[2] + grn[2] * ear) * env[1] + (rk[3] + grn[3] * ear) * env[2];
            } else {
                dum1 = grn[0] * ear;
                dum2 = (grn[1] * env[0] + grn[2] * env[1] + grn[3] * env[2]) * ear;
            }
            int bj = 0;
            while (j >= p->body_start[bj + 1]) bj++;
            for (int b = 0; b < nb; b++)       /* FACTOR_NORMAL: the moving body's panels */
                for (int d = 0; d < 6; d++) {
                    const int c = 6 * b + d;
                    const double complex t1 = (b == bj) ? dum1 * env[d] : 0.0;
                    acc[c] += t1 - dum2 * mx[j + (size_t)c * n];
                }
            for (int h = 0; h < nbeta; h++) {
                double complex dx, dy, dz;
                dinp(&w, xq[0], xq[1], xq[2], p->xw[0], p->xw[1], p->beta[h], &dx, &dy, &dz);
                const double complex dincp = dx * env[0] + dy * env[1] + dz * env[2];
                acc[m + h] += -dum1 * dincp - dum2 * mx[j + (size_t)(m + h) * n];
            }
        }
        /* free term: 2 pi on a hull panel (in its plane and inside its polygon), 4 pi elsewhere */
        int onbody = 0;
        for (int j = 0; j < n && !onbody; j++) {
            const double *en = p->dxyz_p + 6 * j;
            const double tolp = fmax(1.0e-6, 1.0e-3 * p->pnsz[j]), tola = 1.0e-3 * p->pnsz[j] * p->pnsz[j];
            const double *va = p->xyz + 3 * (p->ncon[4 * j] - 1);
            const double dpln = (xp[0] - va[0]) * en[0] + (xp[1] - va[1]) * en[1] + (xp[2] - va[2]) * en[2];
            if (fabs(dpln) > tolp) continue;
            double smin = 1.0e30, smax = -1.0e30;
            for (int k = 0; k < p->ncn[j]; k++) {
                const double *a = p->xyz + 3 * (p->ncon[4 * j + k] - 1);
                const double *b = p->xyz + 3 * (p->ncon[4 * j + (k + 1 < p->ncn[j] ? k + 1 : 0)] - 1);
                const double ev[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
                const double wv[3] = {xp[0] - a[0], xp[1] - a[1], xp[2] - a[2]};
                const double cx = ev[1] * wv[2] - ev[2] * wv[1], cy = ev[2] * wv[0] - ev[0] * wv[2],
                             cz = ev[0] * wv[1] - ev[1] * wv[0];
                const double sv = cx * en[0] + cy * en[1] + cz * en[2];
                smin = fmin(smin, sv);
                smax = fmax(smax, sv);
            }
            if (smin >= -tola || smax <= tola) onbody = 1;
        }
        const double sld = onbody ? 2.0 * PI : 4.0 * PI;
        for (int c = 0; c < ncol; c++) {
            double complex ph = acc[c] / sld;
            if (c >= m && !limit) ph += vinp(&w, xp[0], xp[1], xp[2], p->xw[0], p->xw[1], p->beta[c - m]);
            pot[ip + (size_t)c * nfp] = ph;
            if (prs) prs[ip + (size_t)c * nfp] = limit ? p->rho * ph : I * w1 * p->rho * ph;
            if (elv) elv[ip + (size_t)c * nfp] = limit ? ph : I * w1 / p->g * ph;
        }
        free(acc);
    }
    free(amas); free(bdmp); free(exfc); free(mx);
    return info;
}

/* =============================================================================
 * Motion response -- src/SolveMotionMulti.f90 (LOCALTOGLOBAL :24-40,
 * SolveMotionMulti :46-101): body matrices nbody x 6 x 6 in Fortran order on the
 * block diagonal, the linear solve, then with quadratic damping (max |BQDR| >=
 * 1e-6) the equivalent damping BQDR(i, j) w |x_j| until the summed relative
 * change <= 1e-6 (at most 500 steps here; -1 if not reached). Returns the
 * LAPACK INFO of the last factorisation, or -1.
 * ========================================================================== */
int hmo_solve_motion(int nbody, double w1, const double *amas, const double *bdmp, const double *blnr,
                     const double *bqdr, const double *matx, const double *crs, const double *kstf,
                     const double complex *exfc, double complex *dspl, int *iters)
{
    const int m = 6 * nbody;
    const size_t mm = (size_t)m * m;
    double *g[5];
    const double *loc[5] = {blnr, bqdr, matx, crs, kstf};
    for (int k = 0; k < 5; k++) {
        g[k] = calloc(mm, sizeof(double));
        for (int b = 0; b < nbody; b++)
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++)
                    g[k][(6 * b + i) + (size_t)(6 * b + j) * m] = loc[k][b + (size_t)nbody * (i + 6 * j)];
    }
    double norm = 0.0;
    for (size_t e = 0; e < mm; e++) norm = fmax(norm, fabs(g[1][e]));
    double complex *left = malloc(sizeof(double complex) * mm), *x = calloc(m, sizeof(double complex)),
                   *r = malloc(sizeof(double complex) * m);
    int *ipiv = malloc(sizeof(int) * m), info = 0, it = 0;
    for (int pass = 0;; pass++) {
        for (size_t e = 0; e < mm; e++) {
            const int j = (int)(e / m);
            const double vd = pass ? g[1][e] * w1 * cabs(x[j]) : 0.0;
            left[e] = -w1 * w1 * (g[2][e] + amas[e]) + I * w1 * (bdmp[e] + g[0][e] + vd) + g[3][e] + g[4][e];
        }
        for (int i = 0; i < m; i++) r[i] = -cimag(exfc[i]) - I * creal(exfc[i]);
        info = hmo_zgetrf(m, left, m, ipiv);
        hmo_zgetrs(m, 1, left, m, ipiv, r, m);
        if (pass == 0) {
            memcpy(x, r, sizeof(double complex) * m);
            if (norm < 1.0e-6) break;
            continue;
        }
        it++;
        double rerr = 0.0;
        for (int j = 0; j < m; j++)
            if (cabs(r[j]) > 0.0) rerr += cabs(r[j] - x[j]) / cabs(r[j]);
        memcpy(x, r, sizeof(double complex) * m);
        if (rerr <= 1.0e-6) break;
        if (it >= 500) { info = -1; break; }
    }
    memcpy(dspl, x, sizeof(double complex) * m);
    if (iters) *iters = it;
    for (int k = 0; k < 5; k++) free(g[k]);
    free(left); free(x); free(r); free(ipiv);
    return info;
}
