/*
 * fo_special.c -- special functions of the oracle (FP64).
 *
 * TEST INFRASTRUCTURE ONLY (see fmb_oracle.h).
 *
 * Follows translation.cu:111-212 (cplxExp, gsl_sf_bessel_hl, sphHrmc,
 * rglBasis, sglBasis, aCoeff, bCoeff) and integral.cu:284-331 (dirDev_R).
 * The reference evaluates j_n, y_n and P_n^m through GSL (not in this image);
 * here they are computed by recurrences:
 *   j_n : Miller downward recurrence normalised by sum (2n+1) j_n^2 = 1
 *   y_n : upward recurrence from y_0, y_1
 *   P~_n^m : fully normalised associated Legendre recurrence (no
 *            Condon-Shortley phase, matching sphHrmc's (-1)^m * GSL Plm).
 * Pinned by tests/test_oracle_special.py against scipy golden vectors
 * (tests/golden/special.npz), the Wronskian j_n y_{n-1} - j_{n-1} y_n = 1/x^2,
 * orthonormality of Y_n^m under a high-order sphere rule, and central finite
 * differences for the normal derivative.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fmb_oracle.h"

double fo_acoef(int n, int m) {
  /* translation.cu:189-197 */
  int am = m < 0 ? -m : m;
  if (n < am) return 0.0;
  double nf = n, mf = m;
  return sqrt((nf + 1 + mf) * (nf + 1 - mf) / ((2 * nf + 1) * (2 * nf + 3)));
}

double fo_bcoef(int n, int m) {
  /* translation.cu:199-212 */
  double nf = n, mf = m;
  if (m >= 0 && n >= m)
    return sqrt((nf - mf - 1) * (nf - mf) / ((2 * nf - 1) * (2 * nf + 1)));
  if (-n <= m && m < 0)
    return -sqrt((nf - mf - 1) * (nf - mf) / ((2 * nf - 1) * (2 * nf + 1)));
  return 0.0;
}

void fo_sph_jn(int nmax, double x, double *j) {
  if (nmax < 0) return;
  if (x == 0.0) {
    j[0] = 1.0;
    for (int n = 1; n <= nmax; n++) j[n] = 0.0;
    return;
  }
  int nk = nmax < 1 ? 1 : nmax; /* keep j_1 for the sign test */
  double *f = (double *)malloc(sizeof(double) * (nk + 1));
  double ax = fabs(x);
  double top = (nk > ax ? nk : ax);
  int nst = (int)ceil(top) + 40 + (int)ceil(6.0 * cbrt(top));
  double fp1 = 0.0, fc = 1.0, sum = 0.0;
  for (int n = nst; n >= 0; n--) {
    if (n <= nk) f[n] = fc;
    sum += (2.0 * n + 1.0) * fc * fc;
    if (n == 0) break;
    double fm1 = (2.0 * n + 1.0) / x * fc - fp1;
    fp1 = fc;
    fc = fm1;
    if (fabs(fc) > 1e100) {
      const double s = 1e-100;
      fc *= s;
      fp1 *= s;
      sum *= s * s;
      for (int q = (n - 1 > nk ? nk + 1 : n); q <= nk; q++) f[q] *= s;
    }
  }
  double c = 1.0 / sqrt(sum);
  double j0 = sin(x) / x;
  double j1 = sin(x) / (x * x) - cos(x) / x;
  int neg = (fabs(j0) >= fabs(j1)) ? ((j0 < 0) != (f[0] < 0))
                                   : ((j1 < 0) != (f[1] < 0));
  if (neg) c = -c;
  for (int n = 0; n <= nmax; n++) j[n] = f[n] * c;
  free(f);
}

void fo_sph_yn(int nmax, double x, double *y) {
  if (nmax < 0) return;
  y[0] = -cos(x) / x;
  if (nmax == 0) return;
  y[1] = -cos(x) / (x * x) - sin(x) / x;
  for (int n = 1; n < nmax; n++) {
    y[n + 1] = (2.0 * n + 1.0) / x * y[n] - y[n - 1];
    if (!isfinite(y[n + 1])) {
      for (int q = n + 1; q <= nmax; q++) y[q] = -INFINITY;
      return;
    }
  }
}

/* P[n(n+1)/2+m] = sqrt((2n+1)/4pi (n-m)!/(n+m)!) * Pbar_n^m(ct), m>=0,
 * Pbar without Condon-Shortley phase (sphHrmc, translation.cu:153-169). */
void fo_legendre_table(int nmax, double ct, double st, double *P) {
  if (nmax < 0) return;
  P[0] = 1.0 / sqrt(4.0 * M_PI);
  for (int m = 0; m <= nmax; m++) {
    int imm = m * (m + 1) / 2 + m;
    if (m > 0) {
      int ip = (m - 1) * m / 2 + (m - 1);
      P[imm] = sqrt((2.0 * m + 1.0) / (2.0 * m)) * st * P[ip];
    }
    if (m + 1 <= nmax) {
      int i1 = (m + 1) * (m + 2) / 2 + m;
      P[i1] = sqrt(2.0 * m + 3.0) * ct * P[imm];
    }
    for (int n = m + 2; n <= nmax; n++) {
      double a = sqrt((4.0 * n * n - 1.0) / ((double)n * n - (double)m * m));
      double b = sqrt((((double)(n - 1) * (n - 1)) - (double)m * m) /
                      (4.0 * (n - 1) * (n - 1) - 1.0));
      int i0 = n * (n + 1) / 2 + m;
      int i1 = (n - 1) * n / 2 + m;
      int i2 = (n - 2) * (n - 1) / 2 + m;
      P[i0] = a * (ct * P[i1] - b * P[i2]);
    }
  }
}

static void ylm_cs(int nmax, double ct, double st, double cp, double sp,
                   double *re, double *im) {
  int nt = (nmax + 1) * (nmax + 2) / 2;
  double *P = (double *)malloc(sizeof(double) * nt);
  fo_legendre_table(nmax, ct, st, P);
  double ecr = 1.0, eci = 0.0; /* e^{i m phi} */
  for (int m = 0; m <= nmax; m++) {
    for (int n = m; n <= nmax; n++) {
      double pv = P[n * (n + 1) / 2 + m];
      int ip = n * n + n + m, in = n * n + n - m;
      re[ip] = pv * ecr;
      im[ip] = pv * eci;
      re[in] = pv * ecr;
      im[in] = -pv * eci;
    }
    double t = ecr * cp - eci * sp;
    eci = ecr * sp + eci * cp;
    ecr = t;
  }
  free(P);
}

void fo_ylm(int nmax, double theta, double phi, double *re, double *im) {
  ylm_cs(nmax, cos(theta), sin(theta), cos(phi), sin(phi), re, im);
}

/* cart2sph (structs.cu:48-55): theta = acos(z/r), phi = atan2(y,x) */
static void sph_of(const double *v, double *r, double *ct, double *st,
                   double *cp, double *sp) {
  double rho2 = v[0] * v[0] + v[1] * v[1];
  double rr = sqrt(rho2 + v[2] * v[2]);
  *r = rr;
  if (rr == 0.0) {
    *ct = 1.0;
    *st = 0.0;
    *cp = 1.0;
    *sp = 0.0;
    return;
  }
  double rho = sqrt(rho2);
  *ct = v[2] / rr;
  *st = rho / rr;
  double phi = atan2(v[1], v[0]);
  *cp = cos(phi);
  *sp = sin(phi);
}

void fo_rbasis(int nmax, double k, const double *v, double *re, double *im) {
  double r, ct, st, cp, sp;
  sph_of(v, &r, &ct, &st, &cp, &sp);
  ylm_cs(nmax, ct, st, cp, sp, re, im);
  double *jn = (double *)malloc(sizeof(double) * (nmax + 1));
  fo_sph_jn(nmax, k * r, jn);
  for (int n = 0; n <= nmax; n++)
    for (int m = -n; m <= n; m++) {
      int i = n * n + n + m;
      re[i] *= jn[n];
      im[i] *= jn[n];
    }
  free(jn);
}

void fo_sbasis(int nmax, double k, const double *v, double *re, double *im) {
  double r, ct, st, cp, sp;
  sph_of(v, &r, &ct, &st, &cp, &sp);
  ylm_cs(nmax, ct, st, cp, sp, re, im);
  double *jn = (double *)malloc(sizeof(double) * (nmax + 1));
  double *yn = (double *)malloc(sizeof(double) * (nmax + 1));
  fo_sph_jn(nmax, k * r, jn);
  fo_sph_yn(nmax, k * r, yn);
  for (int n = 0; n <= nmax; n++)
    for (int m = -n; m <= n; m++) {
      int i = n * n + n + m;
      double a = re[i], b = im[i];
      re[i] = jn[n] * a - yn[n] * b;
      im[i] = jn[n] * b + yn[n] * a;
    }
  free(jn);
  free(yn);
}

/* dirDev_R (integral.cu:284-331): normal derivative of R_n^m,
 * dR/dn = k/2 (nx - i ny)[b_{n+1}^{-m-1} R_{n+1}^{m+1} - b_n^m R_{n-1}^{m+1}]
 *       + k/2 (nx + i ny)[b_{n+1}^{m-1} R_{n+1}^{m-1} - b_n^{-m} R_{n-1}^{m-1}]
 *       + k nz [a_{n-1}^m R_{n-1}^m - a_n^m R_{n+1}^m]                     */
void fo_rbasis_dn(int nmax, double k, const double *v, const double *nrm,
                  double *re, double *im) {
  int n1 = nmax + 1;
  int sz = (n1 + 1) * (n1 + 1);
  double *Rr = (double *)malloc(sizeof(double) * sz);
  double *Ri = (double *)malloc(sizeof(double) * sz);
  fo_rbasis(n1, k, v, Rr, Ri);
#define RGET(nn, mm, part)                                                     \
  (((nn) < 0 || abs(mm) > (nn)) ? 0.0                                          \
                                : part[(nn) * (nn) + (nn) + (mm)])
  for (int n = 0; n <= nmax; n++)
    for (int m = -n; m <= n; m++) {
      double t0r = fo_bcoef(n + 1, -m - 1) * RGET(n + 1, m + 1, Rr) -
                   fo_bcoef(n, m) * RGET(n - 1, m + 1, Rr);
      double t0i = fo_bcoef(n + 1, -m - 1) * RGET(n + 1, m + 1, Ri) -
                   fo_bcoef(n, m) * RGET(n - 1, m + 1, Ri);
      double t1r = fo_bcoef(n + 1, m - 1) * RGET(n + 1, m - 1, Rr) -
                   fo_bcoef(n, -m) * RGET(n - 1, m - 1, Rr);
      double t1i = fo_bcoef(n + 1, m - 1) * RGET(n + 1, m - 1, Ri) -
                   fo_bcoef(n, -m) * RGET(n - 1, m - 1, Ri);
      double t2r = fo_acoef(n - 1, m) * RGET(n - 1, m, Rr) -
                   fo_acoef(n, m) * RGET(n + 1, m, Rr);
      double t2i = fo_acoef(n - 1, m) * RGET(n - 1, m, Ri) -
                   fo_acoef(n, m) * RGET(n + 1, m, Ri);
      double hx = 0.5 * k * nrm[0], hy = 0.5 * k * nrm[1];
      /* (hx - i hy) * t0 + (hx + i hy) * t1 + k nz t2 */
      double rr = hx * t0r + hy * t0i + hx * t1r - hy * t1i + k * nrm[2] * t2r;
      double ii = hx * t0i - hy * t0r + hx * t1i + hy * t1r + k * nrm[2] * t2i;
      int idx = n * n + n + m;
      re[idx] = rr;
      im[idx] = ii;
    }
#undef RGET
  free(Rr);
  free(Ri);
}
