/* adjoint.c -- gradient of a field-point objective with respect to the Robin
 * coefficients by the adjoint method.  The reference has no adjoint.
 *
 * Forward problem: A(sigma) x = b, p_q = p_inc(y_q) + sum_j F_qj x_j with
 *   A_ij = delta_ij (1/2 + alpha sigma_i / 2) - K_ij - alpha T_ij
 *          + sigma_j (S_ij + alpha K'_ij),
 *   F_qj = int_j (dG/dn_y - sigma_j G)(y_q, y) dS_y,
 * objective J = sum_r sum_q |p_rq|^2.  A is affine in sigma:
 * dA/dsigma_j = c_j e_j^T with c_j = column j of C = S + alpha K' + alpha I/2,
 * and dF_qj/dsigma_j = -Sf_qj (Sf = the single layer at the field points).
 * With u = conj(p), the holomorphic derivative of sum_q u_q p_q is
 *   w_j = -x_j [ (C^T lambda)_j + (Sf^T u)_j ],   A^T lambda = F^T u,
 * and the gradient of the real J is dJ/dRe sigma_j + i dJ/dIm sigma_j =
 * 2 conj(w_j), summed over right-hand sides.
 *
 * A^T products run the transposed FMM pass (plan.c, rcr.c); C^T lambda runs
 * the same pass with the sigma-derivative P2M basis and near blocks.
 *
 * Vertex gradients (shape.hip): dJ = 2 Re [u^T dF x + lambda^T (db - dA x)]
 * split into per-element terms, with the forward pass's leaf local
 * expansions of x and the transposed pass's leaf multipole expansions of
 * lambda standing in for the far rows and columns of A. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../fmb_internal.h"
#include "../hip/fmb_hip.h"
#include "host.h"

/* inner: NULL for fmb_gmres_transpose, else the adjoint solve refines
 * around GMRES on that plan (fmb_gmres_refined_transpose) */
static int adjoint_run(fmb_plan *pl, fmb_plan *inner, double inner_tol, const fmb_c64 *x,
                       int nrhs, const double *dirs, const double *pts, int npts, int restart,
                       double tol, int maxit, double *J, fmb_c64 *dJ_dsigma, double *dJ_dV,
                       double *dJ_dvn, fmb_gmres_info *info, void *stream) {
  const int shape = dJ_dV || dJ_dvn;
  if (!pl || !pl->ready || !x || nrhs < 1 || nrhs > pl->nrhs_max || !pts ||
      npts < 1 || !(dJ_dsigma || shape) || restart < 1 || !(tol > 0) || maxit < 1)
    return fmb_set_error(FMB_ERR_ARG, "fmb_adjoint_grad: invalid arguments");
  if (shape && !dirs)
    return fmb_set_error(FMB_ERR_ARG,
                         "fmb_adjoint_grad: vertex gradients need the incident directions");
  if (pl->prm.rcr_path == FMB_RCR_VALU)
    return fmb_set_error(FMB_ERR_ARG, "fmb_adjoint_grad: needs rcr_path FMB_RCR_AUTO");
  const fmb_tree *t = pl->tree;
  const int N = t->nelem, L = t->lmax, pL = pl->p[L];
  const size_t nv = (size_t)N * nrhs, np = (size_t)npts * nrhs;
  double *d_pts = NULL, *d_dirs = NULL;
  fmb_c64 *d_p = NULL, *d_kt = NULL, *d_st = NULL, *d_lam = NULL, *d_c = NULL;
  fmb_c64 *h_p = NULL;
  void *d_bs = NULL, *d_nc = NULL, *d_lx = NULL;
  double *d_dv = NULL;
  const size_t lbytes = (size_t)t->nbox[L] * nrhs * pL * pL * sizeof(fmb_c64);
  int rc = fmbh_malloc((void **)&d_pts, sizeof(double) * 3 * (size_t)npts);
  if (!rc) rc = fmbh_h2d(d_pts, pts, sizeof(double) * 3 * (size_t)npts, stream);
  if (!rc && dirs) rc = fmbh_malloc((void **)&d_dirs, sizeof(double) * 3 * nrhs);
  if (!rc && dirs) rc = fmbh_h2d(d_dirs, dirs, sizeof(double) * 3 * nrhs, stream);
  if (!rc) rc = fmbh_malloc((void **)&d_p, sizeof(fmb_c64) * np);
  if (!rc) rc = fmbh_malloc((void **)&d_kt, sizeof(fmb_c64) * nv);
  if (!rc) rc = fmbh_malloc((void **)&d_st, sizeof(fmb_c64) * nv);
  if (!rc) rc = fmbh_malloc((void **)&d_lam, sizeof(fmb_c64) * nv);
  if (!rc) rc = fmbh_malloc((void **)&d_c, sizeof(fmb_c64) * nv);
  if (rc) goto done;
  /* p at the field points, J, u = conj(p) */
  if ((rc = fmbh_field(pl, x, nrhs, d_pts, npts, d_dirs, d_p, stream))) goto done;
  h_p = (fmb_c64 *)malloc(sizeof(fmb_c64) * np);
  if (!h_p) {
    rc = fmb_set_error(FMB_ERR_NOMEM, "fmb_adjoint_grad: host buffer");
    goto done;
  }
  if ((rc = fmbh_d2h(h_p, d_p, sizeof(fmb_c64) * np, stream))) goto done;
  {
    double s = 0.0;
    for (size_t i = 0; i < np; i++) {
      s += h_p[i].re * h_p[i].re + h_p[i].im * h_p[i].im;
      h_p[i].im = -h_p[i].im;
    }
    if (J) *J = s;
  }
  if ((rc = fmbh_h2d(d_p, h_p, sizeof(fmb_c64) * np, stream))) goto done;
  /* leaf local expansions of x from an FP64 pass (vertex gradients) */
  if (shape) {
    if ((rc = fmbh_malloc(&d_lx, lbytes))) goto done;
    if ((rc = fmb_matvec(pl, FMB_FP64, x, d_c, nrhs, stream))) goto done;
    if ((rc = fmbh_d2d(d_lx, pl->ops[0].d_L[L], lbytes, stream))) goto done;
  }
  /* F^T u and Sf^T u */
  if ((rc = fmbh_field_t(pl, d_p, nrhs, d_pts, npts, d_kt, d_st, stream))) goto done;
  /* A^T lambda = F^T u */
  if ((rc = fmbh_memset(d_lam, 0, sizeof(fmb_c64) * nv, stream))) goto done;
  rc = inner ? fmb_gmres_refined_transpose(pl, inner, d_kt, d_lam, nrhs, restart, tol, inner_tol,
                                           maxit, info, stream)
             : fmb_gmres_transpose(pl, d_kt, d_lam, nrhs, restart, tol, maxit, info, stream);
  if (rc) goto done;
  /* C^T lambda in FP64 with the sigma-derivative basis and near blocks */
  if ((rc = fmbh_malloc(&d_bs, sizeof(fmb_c64) * (size_t)N * pL * pL))) goto done;
  if ((rc = fmbh_malloc(&d_nc, sizeof(fmb_c64) * (pl->near_total ? pl->near_total : 1))))
    goto done;
  if ((rc = fmbh_bases_dsigma(pl, 0, d_bs, stream))) goto done;
  if ((rc = fmbh_near_dsigma(pl, 0, d_nc, stream))) goto done;
  if ((rc = fmbx_matvec_t(pl, 0, 3, d_nc, d_bs, d_lam, d_c, nrhs, stream))) goto done;
  if (dJ_dsigma && (rc = fmbh_adjoint_combine(x, d_c, d_st, nrhs, N, dJ_dsigma, stream)))
    goto done;
  /* vertex gradients: the last FP64 transposed pass left lambda's leaf
   * multipole expansions in the FP64 opset */
  if (shape) {
    double *dv = dJ_dV;
    if (!dv) {
      if ((rc = fmbh_malloc((void **)&d_dv, sizeof(double) * 3 * (size_t)t->nvert)))
        goto done;
      dv = d_dv;
    }
    if ((rc = fmbh_shape_grad(pl, x, d_lam, nrhs, d_lx, pl->ops[0].d_M[L], d_p, d_pts, npts,
                              d_dirs, dv, dJ_dvn, stream)))
      goto done;
  }
  rc = fmbh_sync(stream);
done:
  free(h_p);
  fmbh_free(d_pts);
  fmbh_free(d_dirs);
  fmbh_free(d_p);
  fmbh_free(d_kt);
  fmbh_free(d_st);
  fmbh_free(d_lam);
  fmbh_free(d_c);
  fmbh_free(d_bs);
  fmbh_free(d_nc);
  fmbh_free(d_lx);
  fmbh_free(d_dv);
  return rc;
}

int fmb_adjoint_grad(fmb_plan *pl, const fmb_c64 *x, int nrhs, const double *dirs,
                     const double *pts, int npts, int restart, double tol, int maxit,
                     double *J, fmb_c64 *dJ_dsigma, double *dJ_dV, double *dJ_dvn,
                     fmb_gmres_info *info, void *stream) {
  return adjoint_run(pl, NULL, 0.0, x, nrhs, dirs, pts, npts, restart, tol, maxit, J, dJ_dsigma,
                     dJ_dV, dJ_dvn, info, stream);
}

int fmb_adjoint_grad_refined(fmb_plan *pl, fmb_plan *inner, double inner_tol,
                             const fmb_c64 *x, int nrhs, const double *dirs, const double *pts,
                             int npts, int restart, double tol, int maxit, double *J,
                             fmb_c64 *dJ_dsigma, double *dJ_dV, double *dJ_dvn,
                             fmb_gmres_info *info, void *stream) {
  if (!inner) return fmb_set_error(FMB_ERR_ARG, "fmb_adjoint_grad_refined: no inner plan");
  return adjoint_run(pl, inner, inner_tol, x, nrhs, dirs, pts, npts, restart, tol, maxit, J,
                     dJ_dsigma, dJ_dV, dJ_dvn, info, stream);
}
