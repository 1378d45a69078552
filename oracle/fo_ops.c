/*
 * fo_ops.c -- rotate-coaxial-rotate translation operators of the oracle
 * (FP64).
 *
 * TEST INFRASTRUCTURE ONLY (see fmb_oracle.h).
 *
 *   sparse layouts        translation.cu:12-56 (sparseRotSize,
 *                         sparseCoaxTransSize, sparseCoaxMatBlockIdx)
 *   rotation              rotMatsInit translation.cu:1064-1092 (initial
 *                         column H_n^{m',0}), sparseRotMatGen
 *                         translation.cu:1171-1237 (recurrence in m,
 *                         symmetry for m<0, phase e^{i m alpha} e^{-i m' gamma})
 *   coaxial               rr/srCoaxTransMatsInit translation.cu:555-584,
 *                         919-949 (initial column (-1)^n' sqrt(2n'+1) f_n'(kt));
 *                         sparseCoaxTransMatGen translation.cu:678-771
 *                         (sectorial recurrence, advancing in n, |m| blocks)
 *   block products        cuSparseRotVecMul / cuSparseCoaxTransMatVecMul
 *                         translation.cu:1552-1611
 *   RCR composition       transMatsVecsMul_*_rcr translation.cu:1711-1808:
 *                         rotate (alpha,beta,gamma) = (phi,theta,0), coaxial by
 *                         |t|, rotate (0,theta,phi).
 * Where the reference generates in FP32 with (2p-1)^4 scratch, this works in
 * FP64 with O(p^3) scratch.  Pinned by tests/test_oracle_ops.py: unitarity of
 * every rotation block, identity at beta = 0, closed-form Wigner d^1, the
 * expansion of R_n^m about a rotated frame, the (R|R)/(S|R) addition theorems
 * evaluated directly, the reference's own invariants sparse == dense and
 * reduce(p1 -> p2) == generate(p2) (main.c:29-38), and RCR translations of
 * multipole fields checked against direct evaluation.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fmb_oracle.h"

int fo_rot_size(int p) { return p * (4 * p * p - 1) / 3; }
int fo_coax_size(int p) { return p * (2 * p * p + 3 * p + 1) / 6; }
int fo_coax_block(int p, int m) {
  int a = m < 0 ? -m : m;
  return a * (2 * a * a - (3 + 6 * p) * a + 1 + 6 * (p + p * p)) / 6;
}

/* Wigner d^l_{m'm}(beta) for l = l0..L, l0 = max(|m'|,|m|), by the
 * three-term recurrence in the degree (Edmonds 1957, eq. 4.1.23 family),
 * started from the closed form at l0.  Stable for all orders used here,
 * unlike the m-recurrence of the reference (see fo_rot_H_ref). */
static double d_top(int j, int m, double beta) {
  double c = cos(0.5 * beta), s = sin(0.5 * beta);
  double lv = 0.5 * (lgamma(2.0 * j + 1) - lgamma(j + m + 1.0) - lgamma(j - m + 1.0));
  double sg = ((j - m) & 1) ? -1.0 : 1.0;
  if (j + m > 0) {
    if (c == 0.0) return 0.0;
    lv += (j + m) * log(fabs(c));
    if (c < 0 && ((j + m) & 1)) sg = -sg;
  }
  if (j - m > 0) {
    if (s == 0.0) return 0.0;
    lv += (j - m) * log(fabs(s));
    if (s < 0 && ((j - m) & 1)) sg = -sg;
  }
  return sg * exp(lv);
}

static double d_start(int mp, int m, double beta) {
  int j = abs(mp) > abs(m) ? abs(mp) : abs(m);
  if (mp == j) return d_top(j, m, beta);
  if (mp == -j) return ((j + m) & 1 ? -1.0 : 1.0) * d_top(j, -m, beta);
  if (m == j) return ((j - mp) & 1 ? -1.0 : 1.0) * d_top(j, mp, beta);
  return d_top(j, -mp, beta); /* m == -j */
}

/* H_n^{m',m}(beta) for n < p in the sparse layout: block n at n(4n^2-1)/3,
 * element (m'+n, m+n) column-major with leading dimension 2n+1.
 * H_n^{m',m} = eps_{m'} (-1)^m eps_m d^n_{m'm}(beta), eps_m = (-1)^m for
 * m >= 0 and 1 for m < 0: the normalisation of rotMatsInit/sparseRotMatGen
 * (translation.cu:1064-1237), checked against fo_rot_H_ref at small p. */
void fo_rot_H(int p, double beta, double *H) {
  double cb = cos(beta);
  for (int mp = -(p - 1); mp <= p - 1; mp++)
    for (int m = -(p - 1); m <= p - 1; m++) {
      int l0 = abs(mp) > abs(m) ? abs(mp) : abs(m);
      double em = (m < 0 ? ((-m) & 1 ? -1.0 : 1.0) : 1.0);
      double ep = (mp > 0 && (mp & 1)) ? -1.0 : 1.0;
      double f = em * ep;
      double dm2 = 0.0, dm1 = d_start(mp, m, beta);
      for (int n = l0; n < p; n++) {
        double v;
        if (n == l0)
          v = dm1;
        else if (n == 1)
          v = cb; /* only m' = m = 0 reaches here */
        else {
          double a = (2.0 * n - 1.0) * (n * (n - 1.0) * cb - (double)mp * m);
          double b = n * sqrt(fmax(0.0, ((n - 1.0) * (n - 1.0) - (double)mp * mp) *
                                              ((n - 1.0) * (n - 1.0) - (double)m * m)));
          double den = (n - 1.0) * sqrt(((double)n * n - (double)mp * mp) *
                                        ((double)n * n - (double)m * m));
          v = (a * dm1 - b * dm2) / den;
        }
        if (n > l0) {
          dm2 = dm1;
          dm1 = v;
        }
        H[n * (4 * n * n - 1) / 3 + (m + n) * (2 * n + 1) + (mp + n)] = f * v;
      }
    }
}

/* The reference's recurrence in m (sparseRotMatGen, translation.cu:1171-1237)
 * restated in FP64.  It loses about one digit per degree beyond n ~ 15 (block
 * orthogonality 4e-6 at p = 20, O(1) at p = 30), so it serves only as the
 * small-p parity check of fo_rot_H. */
void fo_rot_H_ref(int p, double beta, double *H) {
  int nmax = 2 * p - 2;
  if (nmax < 0) return;
  double cb = cos(beta), sb = sin(beta);
  /* current column m for degrees 0..nmax: col[n*n + n + m'] */
  int sz = (nmax + 1) * (nmax + 1);
  double *col = (double *)calloc(sz, sizeof(double));
  double *nxt = (double *)calloc(sz, sizeof(double));
  int nt = (nmax + 1) * (nmax + 2) / 2;
  double *P = (double *)malloc(sizeof(double) * nt);
  fo_legendre_table(nmax, cb, sb, P);
  for (int n = 0; n <= nmax; n++)
    for (int mp = -n; mp <= n; mp++) {
      int am = mp < 0 ? -mp : mp;
      col[n * n + n + mp] = sqrt(4.0 * M_PI / (2.0 * n + 1.0)) *
                            P[n * (n + 1) / 2 + am];
    }
  free(P);
  /* store m = 0 column */
  for (int n = 0; n < p; n++) {
    int off = n * (4 * n * n - 1) / 3, ld = 2 * n + 1;
    for (int mp = -n; mp <= n; mp++)
      H[off + (0 + n) * ld + (mp + n)] = col[n * n + n + mp];
  }
  for (int m = 0; m + 1 <= p - 1; m++) {
    for (int n = m + 1; n <= nmax - (m + 1); n++) {
      double bd = fo_bcoef(n + 1, m);
      for (int mp = -n; mp <= n; mp++) {
        double t0 = 0.5 * fo_bcoef(n + 1, -mp - 1) * (1.0 - cb);
        double t1 = 0.5 * fo_bcoef(n + 1, mp - 1) * (1.0 + cb);
        double t2 = fo_acoef(n, mp) * sb;
        int b1 = (n + 1) * (n + 1) + (n + 1);
        double v = t0 * col[b1 + mp + 1] - t1 * col[b1 + mp - 1] -
                   t2 * col[b1 + mp];
        nxt[n * n + n + mp] = v / bd;
      }
    }
    /* swap: nxt becomes column m+1 (valid for n >= m+1) */
    double *tmp = col;
    col = nxt;
    nxt = tmp;
    for (int n = m + 1; n < p; n++) {
      int off = n * (4 * n * n - 1) / 3, ld = 2 * n + 1;
      for (int mp = -n; mp <= n; mp++)
        H[off + (m + 1 + n) * ld + (mp + n)] = col[n * n + n + mp];
    }
  }
  /* negative m: H_n^{m',m} = H_n^{-m',-m}  (translation.cu:1206-1219) */
  for (int n = 1; n < p; n++) {
    int off = n * (4 * n * n - 1) / 3, ld = 2 * n + 1;
    for (int m = -n; m < 0; m++)
      for (int mp = -n; mp <= n; mp++)
        H[off + (m + n) * ld + (mp + n)] = H[off + (-m + n) * ld + (-mp + n)];
  }
  free(col);
  free(nxt);
}

void fo_rot_sparse(int p, double alpha, double beta, double gamma,
                   double *re, double *im) {
  int sz = fo_rot_size(p);
  double *H = (double *)malloc(sizeof(double) * (sz > 0 ? sz : 1));
  fo_rot_H(p, beta, H);
  for (int n = 0; n < p; n++) {
    int off = n * (4 * n * n - 1) / 3, ld = 2 * n + 1;
    for (int m = -n; m <= n; m++)
      for (int mp = -n; mp <= n; mp++) {
        double ph = m * alpha - mp * gamma;
        double h = H[off + (m + n) * ld + (mp + n)];
        re[off + (m + n) * ld + (mp + n)] = h * cos(ph);
        im[off + (m + n) * ld + (mp + n)] = h * sin(ph);
      }
  }
  free(H);
}

/* E(m, n', n), m in [0,p), n' in [0, 2p-2], n in [0, p) */
#define EIDX(m, np, n) ((((m) * (2 * p - 1)) + (np)) * p + (n))

void fo_coax_sparse(int kind, int p, double k, double t, double *re,
                    double *im) {
  if (p <= 0) return;
  int nmax = 2 * p - 2;
  size_t sz = (size_t)p * (2 * p - 1) * p;
  fo_cplx *E = (fo_cplx *)calloc(sz, sizeof(fo_cplx));
  double *jn = (double *)malloc(sizeof(double) * (nmax + 1));
  double *yn = (double *)malloc(sizeof(double) * (nmax + 1));
  fo_sph_jn(nmax, k * t, jn);
  fo_sph_yn(nmax, k * t, yn);
  for (int np = 0; np <= nmax; np++) {
    double s = ((np & 1) ? -1.0 : 1.0) * sqrt(2.0 * np + 1.0);
    fo_cplx f = (kind == FO_COAX_SR) ? (jn[np] + I * yn[np]) : jn[np];
    E[EIDX(0, np, 0)] = s * f;
  }
  free(jn);
  free(yn);
  /* sectorial (translation.cu:689-711) */
  for (int m = 0; m + 1 <= p - 1; m++) {
    double den = fo_bcoef(m + 1, -m - 1);
    for (int np = m + 1; np <= nmax - (m + 1); np++) {
      double a0 = fo_bcoef(np, -m - 1) / den;
      double a1 = fo_bcoef(np + 1, m) / den;
      E[EIDX(m + 1, np, m + 1)] =
          a0 * E[EIDX(m, np - 1, m)] - a1 * E[EIDX(m, np + 1, m)];
    }
  }
  /* advancing in n (translation.cu:713-745) */
  for (int m = 0; m <= p - 2; m++)
    for (int n = m; n + 1 <= p - 1; n++) {
      double an = fo_acoef(n, m);
      for (int np = m; np <= nmax - (n + 1); np++) {
        fo_cplx v0 = (np - 1 >= m) ? E[EIDX(m, np - 1, n)] : 0.0;
        fo_cplx v1 = E[EIDX(m, np + 1, n)];
        fo_cplx v2 = (n - 1 >= m) ? E[EIDX(m, np, n - 1)] : 0.0;
        E[EIDX(m, np, n + 1)] = (fo_acoef(np - 1, m) * v0 -
                                 fo_acoef(np, m) * v1 +
                                 fo_acoef(n - 1, m) * v2) /
                                an;
      }
    }
  /* sparse store (translation.cu:760-769).  Only the entries with n' >= n
   * are taken from the recurrence: for n' < n the advancing step subtracts
   * terms (kt)^-2 larger than its result and loses accuracy when kt is
   * small.  They follow from the exact symmetry
   * (E|F)^m_{n'n} = (-1)^{n+n'} (E|F)^m_{nn'} of coaxial coefficients. */
  for (int m = 0; m < p; m++) {
    int off = fo_coax_block(p, m), ld = p - m;
    for (int np = m; np < p; np++)
      for (int n = m; n < p; n++) {
        fo_cplx v = np >= n ? E[EIDX(m, np, n)]
                            : (((n + np) & 1) ? -1.0 : 1.0) * E[EIDX(m, n, np)];
        re[off + (n - m) * ld + (np - m)] = creal(v);
        im[off + (n - m) * ld + (np - m)] = cimag(v);
      }
  }
  free(E);
}
#undef EIDX

void fo_rot_apply(int p, const double *mre, const double *mim,
                  const double *xre, const double *xim, double *yre,
                  double *yim) {
  for (int n = 0; n < p; n++) {
    int off = n * (4 * n * n - 1) / 3, ld = 2 * n + 1;
    for (int mp = -n; mp <= n; mp++) {
      double sr = 0.0, si = 0.0;
      for (int m = -n; m <= n; m++) {
        int e = off + (m + n) * ld + (mp + n);
        int xi = n * n + n + m;
        sr += mre[e] * xre[xi] - mim[e] * xim[xi];
        si += mre[e] * xim[xi] + mim[e] * xre[xi];
      }
      yre[n * n + n + mp] = sr;
      yim[n * n + n + mp] = si;
    }
  }
}

/* y (order pout) = C x (order pin) with C generated at order P >= pin,pout */
void fo_coax_apply(int P, int pin, int pout, const double *mre,
                   const double *mim, const double *xre, const double *xim,
                   double *yre, double *yim) {
  for (int i = 0; i < pout * pout; i++) yre[i] = yim[i] = 0.0;
  int pm = pin < pout ? pin : pout;
  for (int m = -(pm - 1); m <= pm - 1; m++) {
    int am = m < 0 ? -m : m;
    int off = fo_coax_block(P, am), ld = P - am;
    for (int np = am; np < pout; np++) {
      double sr = 0.0, si = 0.0;
      for (int n = am; n < pin; n++) {
        int e = off + (n - am) * ld + (np - am);
        int xi = n * n + n + m;
        sr += mre[e] * xre[xi] - mim[e] * xim[xi];
        si += mre[e] * xim[xi] + mim[e] * xre[xi];
      }
      yre[np * np + np + m] = sr;
      yim[np * np + np + m] = si;
    }
  }
}

/* full RCR translation by vector t (translation.cu:1711-1808) */
void fo_translate(int kind, int pin, int pout, double k, const double *t,
                  const double *xre, const double *xim, double *yre,
                  double *yim) {
  int P = pin > pout ? pin : pout;
  double r = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  double theta = acos(t[2] / r), phi = atan2(t[1], t[0]);
  int rs = fo_rot_size(P), cs = fo_coax_size(P), vs = P * P;
  double *Rr = (double *)malloc(sizeof(double) * rs);
  double *Ri = (double *)malloc(sizeof(double) * rs);
  double *Cr = (double *)malloc(sizeof(double) * cs);
  double *Ci = (double *)malloc(sizeof(double) * cs);
  double *ar = (double *)calloc(vs, sizeof(double));
  double *ai = (double *)calloc(vs, sizeof(double));
  double *br = (double *)calloc(vs, sizeof(double));
  double *bi = (double *)calloc(vs, sizeof(double));
  memcpy(ar, xre, sizeof(double) * pin * pin);
  memcpy(ai, xim, sizeof(double) * pin * pin);
  fo_rot_sparse(P, phi, theta, 0.0, Rr, Ri);
  fo_rot_apply(P, Rr, Ri, ar, ai, br, bi);
  fo_coax_sparse(kind, P, k, r, Cr, Ci);
  fo_coax_apply(P, pin, pout, Cr, Ci, br, bi, ar, ai);
  fo_rot_sparse(pout, 0.0, theta, phi, Rr, Ri);
  fo_rot_apply(pout, Rr, Ri, ar, ai, yre, yim);
  free(Rr);
  free(Ri);
  free(Cr);
  free(Ci);
  free(ar);
  free(ai);
  free(br);
  free(bi);
}
