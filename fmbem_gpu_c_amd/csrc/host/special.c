/* special.c -- host FP64 special functions and quadrature rules:
 * truncNum (octree.c:777-796), Gauss-Legendre nodes (genGaussParams,
 * integral.cu:23-63, by Newton iteration), triangle rules, spherical Bessel
 * functions for the expansion scale factors. */
#include <math.h>
#include <stdlib.h>

#include "../fmb_internal.h"
#include "host.h"

int fmb_trunc_num(double k, double eps, double sigma, double a) {
  double t0 = eps * pow(1.0 - 1.0 / sigma, 1.5);
  double p_lo = 1.0 - log(t0) / log(sigma);
  double ka = k * a;
  double p_hi = ka + pow(3.0 * log(1.0 / eps), 1.5) / 2.0 * pow(ka, 1.0 / 3.0);
  return (int)ceil(pow(pow(p_lo, 4) + pow(p_hi, 4), 0.25));
}

int fmb_rot_size(int p) { return p * (4 * p * p - 1) / 3; }
int fmb_coax_size(int p) { return p * (2 * p * p + 3 * p + 1) / 6; }

void fmbx_gauss_legendre(int n, double *x, double *w) {
  for (int i = 0; i < (n + 1) / 2; i++) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 1.0;
    for (int it = 0; it < 100; it++) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; j++) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      double dz = p1 / pp;
      z -= dz;
      if (fabs(dz) < 1e-16) break;
    }
    double p1 = 1.0, p2 = 0.0;
    for (int j = 1; j <= n; j++) {
      double p3 = p2;
      p2 = p1;
      p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
    }
    pp = n * (z * p1 - p2) / (z * z - 1.0);
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
}

/* rule -> nq rows {b0, b1, b2, w}, weights summing to 1 */
int fmbx_tri_rule(int rule, int order, double **out) {
  if (rule == FMB_RULE_GAUSS7) {
    const double s15 = sqrt(15.0);
    const double a1 = (6.0 - s15) / 21.0, b1 = (9.0 + 2.0 * s15) / 21.0;
    const double a2 = (6.0 + s15) / 21.0, b2 = (9.0 - 2.0 * s15) / 21.0;
    const double w1 = (155.0 - s15) / 1200.0, w2 = (155.0 + s15) / 1200.0;
    double T[7][4] = {{1.0 / 3, 1.0 / 3, 1.0 / 3, 9.0 / 40.0},
                      {b1, a1, a1, w1}, {a1, b1, a1, w1}, {a1, a1, b1, w1},
                      {b2, a2, a2, w2}, {a2, b2, a2, w2}, {a2, a2, b2, w2}};
    double *q = (double *)malloc(sizeof(double) * 28);
    for (int i = 0; i < 7; i++)
      for (int c = 0; c < 4; c++) q[4 * i + c] = T[i][c];
    *out = q;
    return 7;
  }
  if (order < 1 || order > 64) return -1;
  /* collapsed rule (integral.cu:80-109) */
  double *gx = (double *)malloc(sizeof(double) * order);
  double *gw = (double *)malloc(sizeof(double) * order);
  fmbx_gauss_legendre(order, gx, gw);
  double *q = (double *)malloc(sizeof(double) * 4 * order * order);
  int k = 0;
  for (int n = 0; n < order; n++) {
    double th = 0.5 + 0.5 * gx[n];
    for (int m = 0; m < order; m++) {
      double rho = 0.5 + 0.5 * gx[m];
      double x1 = rho * (1.0 - th), x2 = rho - x1;
      q[4 * k] = x1;
      q[4 * k + 1] = x2;
      q[4 * k + 2] = 1.0 - x1 - x2;
      q[4 * k + 3] = 0.5 * gw[n] * gw[m] * rho;
      k++;
    }
  }
  free(gx);
  free(gw);
  *out = q;
  return k;
}

/* |h_n(x)| for n = 0..nmax by upward recurrence of y_n and Miller j_n */
void fmbx_abs_hn(int nmax, double x, double *out) {
  int nk = nmax < 1 ? 1 : nmax;
  double top = nk > x ? nk : x;
  int nst = (int)ceil(top) + 40 + (int)ceil(6.0 * cbrt(top));
  double *j = (double *)malloc(sizeof(double) * (nk + 1));
  double fp1 = 0, fc = 1, sum = 0;
  for (int n = nst; n >= 0; n--) {
    if (n <= nk) j[n] = fc;
    sum += (2.0 * n + 1.0) * fc * fc;
    if (n == 0) break;
    double fm = (2.0 * n + 1.0) / x * fc - fp1;
    fp1 = fc;
    fc = fm;
    if (fabs(fc) > 1e100) {
      fc *= 1e-100;
      fp1 *= 1e-100;
      sum *= 1e-200;
      for (int q = (n - 1 > nk ? nk + 1 : n); q <= nk; q++) j[q] *= 1e-100;
    }
  }
  double c = 1.0 / sqrt(sum);
  double y0 = -cos(x) / x, y1 = -cos(x) / (x * x) - sin(x) / x;
  for (int n = 0; n <= nmax; n++) {
    double y;
    if (n == 0)
      y = y0;
    else if (n == 1)
      y = y1;
    else {
      y = (2.0 * n - 1.0) / x * y1 - y0;
      y0 = y1;
      y1 = y;
    }
    double jj = j[n] * c;
    out[n] = isfinite(y) ? sqrt(jj * jj + y * y) : INFINITY;
  }
  free(j);
}
