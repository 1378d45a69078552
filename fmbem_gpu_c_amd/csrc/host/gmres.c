/* gmres.c -- restarted GMRES (Saad & Schultz 1986) on the device, batched
 * over right-hand sides in lock step.  The reference has no solver.
 *
 * Arnoldi vectors, inner products and the least-squares update are FP64
 * (classical Gram-Schmidt with one reorthogonalisation pass; complex Givens
 * rotations on the host).  Every restart recomputes the true residual
 * b - A x with the FP64 FMM pass; on an FMB_MIXED plan the Arnoldi matvecs use
 * the FP32 pass, so each restart cycle is one step of iterative refinement
 * on the FP64 residual. */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../fmb_internal.h"
#include "../hip/fmb_hip.h"
#include "host.h"

typedef double _Complex zc;

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static zc toz(fmb_c64 v) { return v.re + I * v.im; }
static fmb_c64 fromz(zc z) {
  fmb_c64 v = {creal(z), cimag(z)};
  return v;
}

/* y = A x, or A^T x for the transposed system */
static int apply(fmb_plan *pl, int trans, int prec, const fmb_c64 *x, fmb_c64 *y,
                 int nrhs, void *stream) {
  if (!trans) return fmb_matvec(pl, prec, x, y, nrhs, stream);
  int op = (prec == FMB_MIXED && pl->nops == 2) ? 1 : 0;
  return fmbx_matvec_t(pl, op, 3, NULL, NULL, x, y, nrhs, stream);
}

static int gmres_run(fmb_plan *pl, int trans, const fmb_c64 *b, fmb_c64 *x,
                     int nrhs, int restart, double tol, int maxit,
                     fmb_gmres_info *info, void *stream) {
  if (!pl || !pl->ready || !b || !x || nrhs < 1 || nrhs > pl->nrhs_max || restart < 1 ||
      !(tol > 0) || maxit < 1)
    return fmb_set_error(FMB_ERR_ARG, "fmb_gmres: invalid arguments");
  if (trans && pl->prm.rcr_path == FMB_RCR_VALU)
    return fmb_set_error(FMB_ERR_ARG, "transposed solve needs rcr_path FMB_RCR_AUTO");
  double t0 = now_s();
  const int N = pl->tree->nelem, m = restart;
  const int work = pl->work_prec;
  size_t vbytes = sizeof(fmb_c64) * (size_t)N * nrhs;
  int rc;
  fmb_c64 *dV = NULL, *dw = NULL, *dr = NULL, *dh = NULL, *dh2 = NULL, *dy = NULL;
  double *dn = NULL;
  zc *Hh = NULL, *cs = NULL, *sn = NULL, *g = NULL, *yv = NULL;
  fmb_c64 *hh = NULL, *hh2 = NULL;
  double *bn = NULL, *rn = NULL, *hn = NULL;
  int *active = NULL;
  int iters = 0, mv64 = 0, mvw = 0, restarts = 0, conv = 0;
  if ((rc = fmbh_malloc((void **)&dV, vbytes * (m + 1)))) goto done;
  if ((rc = fmbh_malloc((void **)&dw, vbytes))) goto done;
  if ((rc = fmbh_malloc((void **)&dr, vbytes))) goto done;
  if ((rc = fmbh_malloc((void **)&dh, sizeof(fmb_c64) * (m + 1) * nrhs))) goto done;
  if ((rc = fmbh_malloc((void **)&dh2, sizeof(fmb_c64) * (m + 1) * nrhs))) goto done;
  if ((rc = fmbh_malloc((void **)&dy, sizeof(fmb_c64) * (m + 1) * nrhs))) goto done;
  if ((rc = fmbh_malloc((void **)&dn, sizeof(double) * nrhs))) goto done;
  Hh = (zc *)calloc((size_t)(m + 1) * m * nrhs, sizeof(zc));
  cs = (zc *)calloc((size_t)m * nrhs, sizeof(zc));
  sn = (zc *)calloc((size_t)m * nrhs, sizeof(zc));
  g = (zc *)calloc((size_t)(m + 1) * nrhs, sizeof(zc));
  yv = (zc *)calloc((size_t)(m + 1) * nrhs, sizeof(zc));
  hh = (fmb_c64 *)malloc(sizeof(fmb_c64) * (m + 1) * nrhs);
  hh2 = (fmb_c64 *)malloc(sizeof(fmb_c64) * (m + 1) * nrhs);
  bn = (double *)malloc(sizeof(double) * nrhs);
  rn = (double *)malloc(sizeof(double) * nrhs);
  hn = (double *)malloc(sizeof(double) * nrhs);
  active = (int *)malloc(sizeof(int) * nrhs);
#define HH(r, i, j) Hh[((size_t)(r) * m + (j)) * (m + 1) + (i)]
  if ((rc = fmbh_znorms(b, nrhs, N, dn, stream))) goto done;
  if ((rc = fmbh_d2h(bn, dn, sizeof(double) * nrhs, stream))) goto done;
  for (int r = 0; r < nrhs; r++)
    if (bn[r] == 0.0) bn[r] = 1.0;
  while (1) {
    /* true residual r = b - A x in FP64 */
    if ((rc = apply(pl, trans, FMB_FP64, x, dw, nrhs, stream))) goto done;
    mv64++;
    {
      fmb_c64 *one = hh, *mone = hh2;
      for (int r = 0; r < nrhs; r++) {
        one[r].re = 1.0;
        one[r].im = 0.0;
        mone[r].re = -1.0;
        mone[r].im = 0.0;
      }
      if ((rc = fmbh_h2d(dh, one, sizeof(fmb_c64) * nrhs, stream))) goto done;
      if ((rc = fmbh_h2d(dh2, mone, sizeof(fmb_c64) * nrhs, stream))) goto done;
      if ((rc = fmbh_d2d(dr, b, vbytes, stream))) goto done;
      /* r = 1*b + (-1)*w  ->  zaxpby(a=-1, x=w, b=1, y=r) */
      if ((rc = fmbh_zaxpby(dh2, dw, dh, dr, nrhs, N, stream))) goto done;
    }
    if ((rc = fmbh_znorms(dr, nrhs, N, dn, stream))) goto done;
    if ((rc = fmbh_d2h(rn, dn, sizeof(double) * nrhs, stream))) goto done;
    conv = 1;
    for (int r = 0; r < nrhs; r++) {
      active[r] = rn[r] / bn[r] > tol;
      if (active[r]) conv = 0;
    }
    if (conv || iters >= maxit) break;
    if (restarts > 0 || iters > 0) restarts++;
    /* V_0 = r / |r| */
    for (int r = 0; r < nrhs; r++) {
      double s = rn[r] > 0 ? 1.0 / rn[r] : 0.0;
      hh[r].re = s;
      hh[r].im = 0.0;
      g[(size_t)r * (m + 1)] = rn[r];
      for (int i = 1; i <= m; i++) g[(size_t)r * (m + 1) + i] = 0.0;
    }
    if ((rc = fmbh_h2d(dh, hh, sizeof(fmb_c64) * nrhs, stream))) goto done;
    if ((rc = fmbh_zaxpby(dh, dr, NULL, dV, nrhs, N, stream))) goto done;
    int j;
    for (j = 0; j < m && iters < maxit; j++) {
      fmb_c64 *Vj = dV + (size_t)j * N * nrhs;
      if ((rc = apply(pl, trans, work, Vj, dw, nrhs, stream))) goto done;
      if (work == FMB_MIXED)
        mvw++;
      else
        mv64++;
      iters++;
      /* two passes of classical Gram-Schmidt */
      for (int pass = 0; pass < 2; pass++) {
        fmb_c64 *hbuf = pass ? hh2 : hh;
        if ((rc = fmbh_zdots(dV, j + 1, dw, nrhs, N, dh, stream))) goto done;
        if ((rc = fmbh_zgemv_sub(dV, j + 1, dh, dw, nrhs, N, stream))) goto done;
        if ((rc = fmbh_d2h(hbuf, dh, sizeof(fmb_c64) * (j + 1) * nrhs, stream)))
          goto done;
      }
      if ((rc = fmbh_znorms(dw, nrhs, N, dn, stream))) goto done;
      if ((rc = fmbh_d2h(hn, dn, sizeof(double) * nrhs, stream))) goto done;
      int all_done = 1;
      for (int r = 0; r < nrhs; r++) {
        for (int i = 0; i <= j; i++)
          HH(r, i, j) = toz(hh[(size_t)i * nrhs + r]) + toz(hh2[(size_t)i * nrhs + r]);
        HH(r, j + 1, j) = hn[r];
        double s = hn[r] > 0 ? 1.0 / hn[r] : 0.0;
        hh[r].re = s;
        hh[r].im = 0.0;
        /* apply previous rotations and form the new one */
        for (int i = 0; i < j; i++) {
          zc h0 = HH(r, i, j), h1 = HH(r, i + 1, j);
          zc c = cs[(size_t)r * m + i], sv = sn[(size_t)r * m + i];
          HH(r, i, j) = conj(c) * h0 + conj(sv) * h1;
          HH(r, i + 1, j) = -sv * h0 + c * h1;
        }
        zc a = HH(r, j, j), bb = HH(r, j + 1, j);
        double den = sqrt(creal(a * conj(a)) + creal(bb * conj(bb)));
        zc c = den > 0 ? a / den : 1.0, sv = den > 0 ? bb / den : 0.0;
        cs[(size_t)r * m + j] = c;
        sn[(size_t)r * m + j] = sv;
        HH(r, j, j) = den;
        HH(r, j + 1, j) = 0.0;
        zc gj = g[(size_t)r * (m + 1) + j];
        g[(size_t)r * (m + 1) + j + 1] = -sv * gj;
        g[(size_t)r * (m + 1) + j] = conj(c) * gj;
        if (active[r] && cabs(g[(size_t)r * (m + 1) + j + 1]) / bn[r] > tol)
          all_done = 0;
      }
      if ((rc = fmbh_h2d(dh, hh, sizeof(fmb_c64) * nrhs, stream))) goto done;
      if ((rc = fmbh_zaxpby(dh, dw, NULL, dV + (size_t)(j + 1) * N * nrhs, nrhs,
                            N, stream)))
        goto done;
      if (all_done) {
        j++;
        break;
      }
    }
    /* x += V y, y from the triangular solve (inactive rhs get y = 0) */
    for (int r = 0; r < nrhs; r++) {
      for (int q = j - 1; q >= 0; q--) {
        zc s = g[(size_t)r * (m + 1) + q];
        for (int c2 = q + 1; c2 < j; c2++) s -= HH(r, q, c2) * yv[(size_t)r * (m + 1) + c2];
        zc d = HH(r, q, q);
        yv[(size_t)r * (m + 1) + q] = (active[r] && cabs(d) > 0) ? s / d : 0.0;
      }
      for (int q = 0; q < j; q++)
        hh[(size_t)q * nrhs + r] = fromz(yv[(size_t)r * (m + 1) + q]);
    }
    if ((rc = fmbh_h2d(dy, hh, sizeof(fmb_c64) * j * nrhs, stream))) goto done;
    if ((rc = fmbh_zgemv_add(dV, j, dy, x, nrhs, N, stream))) goto done;
  }
  rc = conv ? FMB_OK : fmb_set_error(FMB_ERR_NOCONV,
                                     "GMRES: tolerance not reached in %d steps",
                                     iters);
  if (info) {
    memset(info, 0, sizeof(*info));
    info->iters = iters;
    info->matvecs_fp64 = mv64;
    info->matvecs_work = mvw;
    info->restarts = restarts;
    for (int r = 0; r < nrhs && r < 64; r++) info->relres[r] = rn[r] / bn[r];
  }
#undef HH
done:
  if (info) info->seconds = now_s() - t0;
  fmbh_free(dV);
  fmbh_free(dw);
  fmbh_free(dr);
  fmbh_free(dh);
  fmbh_free(dh2);
  fmbh_free(dy);
  fmbh_free(dn);
  free(Hh);
  free(cs);
  free(sn);
  free(g);
  free(yv);
  free(hh);
  free(hh2);
  free(bn);
  free(rn);
  free(hn);
  free(active);
  return rc;
}

int fmb_gmres(fmb_plan *pl, const fmb_c64 *b, fmb_c64 *x, int nrhs,
              int restart, double tol, int maxit, fmb_gmres_info *info,
              void *stream) {
  return gmres_run(pl, 0, b, x, nrhs, restart, tol, maxit, info, stream);
}

int fmb_gmres_transpose(fmb_plan *pl, const fmb_c64 *b, fmb_c64 *x, int nrhs,
                        int restart, double tol, int maxit, fmb_gmres_info *info,
                        void *stream) {
  return gmres_run(pl, 1, b, x, nrhs, restart, tol, maxit, info, stream);
}
