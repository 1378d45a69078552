/*
 * fo_quad.c -- quadrature rules, Green's-function kernels and panel
 * integrals of the oracle (FP64).
 *
 * TEST INFRASTRUCTURE ONLY (see fmb_oracle.h).
 *
 * Follows integral.cu:
 *   genGaussParams       integral.cu:23-63   (Gauss-Legendre; Newton iteration
 *                                            here instead of GSL eigen-solve)
 *   green, pGp1n, pGp2n, p2Gp1np2n          integral.cu:72-78, 391-415, 602-622
 *   triElemIntegral_*_nsgl (collapsed INTORDER^2 rule)
 *                                           integral.cu:80-139, 417-449,
 *                                           510-541, 624-657
 *   triElemIntegral_G_sgl (3 sub-triangles at the centroid)
 *                                           integral.cu:172-249
 *   triElemIntegral_p2Gp1np2n_sgl / theta2rho
 *                                           integral.cu:666-751
 * The self-panel integrals used by the solver (fo_panel_self) reduce the
 * centroid-collocated weakly singular and hypersingular integrals on a flat
 * triangle to contour integrals over its three edges:
 *   int_T G dS       = (1/4pi) oint (e^{ik rho}-1)/(ik) dtheta
 *   f.p. int_T d2G   = -(1/4pi) [ oint e^{ik rho}/rho dtheta - 2 pi i k ]
 *                                           (the reference formula, integral.cu:731-751)
 * with oint dtheta/rho evaluated in closed form per edge.  Pinned by
 * tests/test_oracle_quad.py: k=0 closed forms (disk limits, Gauss
 * solid-angle identity), convergence of the reference-style rules
 * (fo_panel_hyper_ref, collapsed sub-triangle rule) towards the contour
 * values, and rule exactness on polynomials.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fmb_oracle.h"

int fo_gauss_legendre(int n, double *x, double *w) {
  if (n < 1) return -1;
  for (int i = 0; i < (n + 1) / 2; i++) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5));
    double pp = 0.0;
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
    {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; j++) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
    }
    /* ascending order */
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
    w[n - 1 - i] = w[i];
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
  return 0;
}

/* Triangle rules as barycentric weights (b0,b1,b2) on (node0,node1,node2),
 * weights normalised to sum 1 (multiply by the panel area). */
int fo_tri_rule(int rule, int order, double *bary, double *w) {
  if (rule == FO_RULE_GAUSS7) {
    const double s15 = sqrt(15.0);
    const double a1 = (6.0 - s15) / 21.0, b1 = (9.0 + 2.0 * s15) / 21.0;
    const double a2 = (6.0 + s15) / 21.0, b2 = (9.0 - 2.0 * s15) / 21.0;
    const double w1 = (155.0 - s15) / 1200.0, w2 = (155.0 + s15) / 1200.0;
    double B[7][3] = {{1.0 / 3, 1.0 / 3, 1.0 / 3}, {b1, a1, a1}, {a1, b1, a1},
                      {a1, a1, b1},                {b2, a2, a2}, {a2, b2, a2},
                      {a2, a2, b2}};
    double W[7] = {9.0 / 40.0, w1, w1, w1, w2, w2, w2};
    for (int q = 0; q < 7; q++) {
      for (int c = 0; c < 3; c++) bary[3 * q + c] = B[q][c];
      w[q] = W[q];
    }
    return 7;
  }
  /* collapsed rule, integral.cu:80-109: rho = 1/2+eta1/2, theta = 1/2+eta2/2,
   * xi1 = rho(1-theta), xi2 = rho theta, xi3 = 1-rho, weight 1/4 wn wm rho J
   * with J = 2*area */
  double *gx = (double *)malloc(sizeof(double) * order);
  double *gw = (double *)malloc(sizeof(double) * order);
  fo_gauss_legendre(order, gx, gw);
  int q = 0;
  for (int n = 0; n < order; n++) {
    double theta = 0.5 + 0.5 * gx[n];
    for (int m = 0; m < order; m++) {
      double rho = 0.5 + 0.5 * gx[m];
      double xi1 = rho * (1.0 - theta);
      double xi2 = rho - xi1;
      double xi3 = 1.0 - xi1 - xi2;
      bary[3 * q + 0] = xi1;
      bary[3 * q + 1] = xi2;
      bary[3 * q + 2] = xi3;
      w[q] = 0.5 * gw[n] * gw[m] * rho;
      q++;
    }
  }
  free(gx);
  free(gw);
  return q;
}

/* kernels at target x (normal nx) from source y (normal ny):
 * [0] G, [1] dG/dn_y, [2] dG/dn_x, [3] d2G/dn_x dn_y   (integral.cu:72-78,
 * 391-415, 602-622) */
void fo_kernels(double k, const double *x, const double *nx, const double *y,
                const double *ny, double *re4, double *im4) {
  double d[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
  double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  fo_cplx e = cexp(I * k * r);
  fo_cplx g = e / (4.0 * M_PI * r);
  double rnx = (d[0] * nx[0] + d[1] * nx[1] + d[2] * nx[2]) / r;  /* prpn1 */
  double rny = -(d[0] * ny[0] + d[1] * ny[1] + d[2] * ny[2]) / r; /* prpn2 */
  fo_cplx gp = g * (-1.0 / r + I * k);
  fo_cplx dny = gp * rny;
  fo_cplx dnx = gp * rnx;
  double nn = nx[0] * ny[0] + nx[1] * ny[1] + nx[2] * ny[2];
  fo_cplx h = e / (4.0 * M_PI * r * r * r) *
              ((3.0 - k * k * r * r - 3.0 * I * k * r) * rnx * rny +
               (1.0 - I * k * r) * nn);
  fo_cplx out[4] = {g, dny, dnx, h};
  for (int i = 0; i < 4; i++) {
    re4[i] = creal(out[i]);
    im4[i] = cimag(out[i]);
  }
}

static void tri_geom(const double *tri, double *nrm, double *area) {
  /* normal (n0-n2)x(n1-n2), integral.cu:420-421 */
  double a[3], b[3], c[3];
  for (int i = 0; i < 3; i++) {
    a[i] = tri[i] - tri[6 + i];
    b[i] = tri[3 + i] - tri[6 + i];
  }
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
  double J = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  for (int i = 0; i < 3; i++) nrm[i] = c[i] / J;
  *area = 0.5 * J;
}

void fo_panel_regular(double k, const double *tri, const double *x,
                      const double *nx, int rule, int order, double *re4,
                      double *im4) {
  double bary[3 * 64 * 64], w[64 * 64];
  if (rule == FO_RULE_COLLAPSED && order > 64) order = 64;
  int nq = fo_tri_rule(rule, order, bary, w);
  double ny[3], area;
  tri_geom(tri, ny, &area);
  for (int i = 0; i < 4; i++) re4[i] = im4[i] = 0.0;
  for (int q = 0; q < nq; q++) {
    double y[3];
    for (int c = 0; c < 3; c++)
      y[c] = bary[3 * q] * tri[c] + bary[3 * q + 1] * tri[3 + c] +
             bary[3 * q + 2] * tri[6 + c];
    double kr[4], ki[4];
    fo_kernels(k, x, nx, y, ny, kr, ki);
    for (int i = 0; i < 4; i++) {
      re4[i] += w[q] * area * kr[i];
      im4[i] += w[q] * area * ki[i];
    }
  }
}

/* Analytic-contour self-panel terms at the centroid.
 * out: [0] int G, [1] int dG/dn_y (=0), [2] int dG/dn_x (=0), [3] f.p. d2G */
void fo_panel_self(double k, const double *tri, int nedge, double *re4,
                   double *im4) {
  double c[3];
  for (int i = 0; i < 3; i++) c[i] = (tri[i] + tri[3 + i] + tri[6 + i]) / 3.0;
  double gx[128], gw[128];
  if (nedge > 128) nedge = 128;
  fo_gauss_legendre(nedge, gx, gw);
  fo_cplx sS = 0.0, sT = 0.0;
  double sinv = 0.0;
  for (int e = 0; e < 3; e++) {
    const double *P = tri + 3 * e, *Q = tri + 3 * ((e + 1) % 3);
    double ev[3], L = 0.0;
    for (int i = 0; i < 3; i++) {
      ev[i] = Q[i] - P[i];
      L += ev[i] * ev[i];
    }
    L = sqrt(L);
    for (int i = 0; i < 3; i++) ev[i] /= L;
    double cp[3], tp = 0.0;
    for (int i = 0; i < 3; i++) cp[i] = c[i] - P[i];
    for (int i = 0; i < 3; i++) tp += cp[i] * ev[i];
    double F[3], h = 0.0;
    for (int i = 0; i < 3; i++) {
      F[i] = P[i] + tp * ev[i];
      h += (c[i] - F[i]) * (c[i] - F[i]);
    }
    h = sqrt(h);
    double sP = -tp, sQ = L - tp;
    double thP = atan2(sP, h), thQ = atan2(sQ, h);
    sinv += (sin(thQ) - sin(thP)) / h;
    double half = 0.5 * (thQ - thP), mid = 0.5 * (thQ + thP);
    for (int q = 0; q < nedge; q++) {
      double th = mid + half * gx[q];
      double rho = h / cos(th);
      double wq = half * gw[q];
      /* (e^{ik rho}-1)/(ik) = 2 sin(k rho/2) e^{ik rho/2}/k */
      fo_cplx gs;
      if (k == 0.0)
        gs = rho;
      else
        gs = 2.0 * sin(0.5 * k * rho) * cexp(0.5 * I * k * rho) / k;
      sS += wq * gs;
      /* (e^{ik rho}-1)/rho = i k gs / rho */
      sT += wq * (I * k * gs / rho);
    }
  }
  fo_cplx S = sS / (4.0 * M_PI);
  fo_cplx T = -(sT + sinv - 2.0 * M_PI * I * k) / (4.0 * M_PI);
  re4[0] = creal(S);
  im4[0] = cimag(S);
  re4[1] = im4[1] = re4[2] = im4[2] = 0.0;
  re4[3] = creal(T);
  im4[3] = cimag(T);
}

/* Reference-style hypersingular self term: Gauss rule in theta over
 * [0, 2pi] with the centroid-to-edge distance rho(theta)
 * (triElemIntegral_p2Gp1np2n_sgl + theta2rho, integral.cu:666-751),
 * restated in FP64 without the tan/tanf mix. */
void fo_panel_hyper_ref(double k, const double *tri, int order, double *re,
                        double *im) {
  double c[3];
  for (int i = 0; i < 3; i++) c[i] = (tri[i] + tri[3 + i] + tri[6 + i]) / 3.0;
  /* local frame: polar axis from centroid to node 0 */
  double d[3][3], dist[3];
  for (int v = 0; v < 3; v++) {
    dist[v] = 0.0;
    for (int i = 0; i < 3; i++) {
      d[v][i] = tri[3 * v + i] - c[i];
      dist[v] += d[v][i] * d[v][i];
    }
    dist[v] = sqrt(dist[v]);
  }
  double e1[3], nrm[3], area, e2[3];
  for (int i = 0; i < 3; i++) e1[i] = d[0][i] / dist[0];
  tri_geom(tri, nrm, &area);
  e2[0] = nrm[1] * e1[2] - nrm[2] * e1[1];
  e2[1] = nrm[2] * e1[0] - nrm[0] * e1[2];
  e2[2] = nrm[0] * e1[1] - nrm[1] * e1[0];
  double px[3], py[3], ang[3];
  for (int v = 0; v < 3; v++) {
    px[v] = d[v][0] * e1[0] + d[v][1] * e1[1] + d[v][2] * e1[2];
    py[v] = d[v][0] * e2[0] + d[v][1] * e2[1] + d[v][2] * e2[2];
    ang[v] = atan2(py[v], px[v]);
    if (ang[v] < 0) ang[v] += 2.0 * M_PI;
  }
  double *gx = (double *)malloc(sizeof(double) * order);
  double *gw = (double *)malloc(sizeof(double) * order);
  fo_gauss_legendre(order, gx, gw);
  fo_cplx sum = 0.0;
  for (int q = 0; q < order; q++) {
    double th = M_PI * gx[q] + M_PI;
    /* which edge does the ray at angle th hit */
    double ux = cos(th), uy = sin(th), rho = 0.0;
    for (int e = 0; e < 3; e++) {
      int a = e, b = (e + 1) % 3;
      double ex = px[b] - px[a], ey = py[b] - py[a];
      double den = ux * ey - uy * ex;
      if (fabs(den) < 1e-300) continue;
      double t = (px[a] * ey - py[a] * ex) / den;
      double s = (px[a] * uy - py[a] * ux) / den;
      if (t > 0 && s >= -1e-14 && s <= 1 + 1e-14) {
        rho = t;
        break;
      }
    }
    sum += gw[q] * cexp(I * k * rho) / rho;
  }
  (void)ang;
  sum *= M_PI;
  fo_cplx T = -(sum - 2.0 * M_PI * I * k) / (4.0 * M_PI);
  *re = creal(T);
  *im = cimag(T);
  free(gx);
  free(gw);
}
