/* gmres.c -- restarted GMRES (Saad & Schultz 1986) on the device, batched
 * in lock step over right-hand sides and, with fmb_gmres_batch, over plans
 * (frequencies) that share the element count.  The reference has no solver.
 *
 * A batch applies every plan per Arnoldi step; with graph mode the plans'
 * FMM passes are captured once per precision into one hipGraph (one branch
 * per plan, forked streams) between fixed staging vectors.  A plan retires
 * as soon as its estimated residuals meet the tolerance (its columns of x
 * updated at once and its FP64 true residual checked) or its true residual
 * does at a restart; retired plans leave the pass and the graph is captured
 * again.
 *
 * Arnoldi vectors, inner products and the least-squares update are FP64
 * (classical Gram-Schmidt with one reorthogonalisation pass; complex Givens
 * rotations on the host).  Every restart recomputes the true residual
 * b - A x with the FP64 FMM pass; on an FMB_MIXED plan the Arnoldi matvecs use
 * the FP32 pass, so each restart cycle is one step of iterative refinement
 * on the FP64 residual.  Their Arnoldi cycles stop when the estimated
 * residual reaches half the tolerance: the FP32 pass leaves the true residual
 * a few per cent above the estimate, so a stop right at the tolerance could
 * fail the FP64 check and pay a second FP64 matvec and a restart. */
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

/* The system: plan p owns columns [off[p], off[p+1]) of every vector. */
typedef struct {
  fmb_plan *const *pl;
  const int *nr;
  int np, trans, N, graph;
  int *off;
  int *live;          /* plans in the captured graphs                  */
  void *exec[2];      /* graph per precision (0 FP64, 1 working)       */
  void *cs, **bs, *ev_fork, **ev_join;
  fmb_c64 *xs, *ys;   /* staging input; output (= the work vector)     */
} sys_t;

static int plan_apply(fmb_plan *pl, int trans, int prec, const fmb_c64 *x, fmb_c64 *y,
                      int nrhs, void *stream) {
  if (!trans) return fmb_matvec(pl, prec, x, y, nrhs, stream);
  int op = (prec == FMB_MIXED && pl->nops == 2) ? 1 : 0;
  return fmbx_matvec_t(pl, op, 3, NULL, NULL, x, y, nrhs, stream);
}

static void graphs_drop(sys_t *S) {
  for (int i = 0; i < 2; i++) {
    fmbh_graph_destroy(S->exec[i]);
    S->exec[i] = NULL;
  }
}

/* one branch per live plan from xs to ys, forked from the capture stream */
static int graph_capture(sys_t *S, int prec, void **exec) {
  int rc = fmbh_capture_begin(S->cs), rc2;
  if (!rc) rc = fmbh_event_record(S->ev_fork, S->cs);
  for (int p = 0; p < S->np && !rc; p++) {
    if (!S->live[p]) continue;
    fmb_plan *pl = S->pl[p];
    int tm = pl->timing;
    pl->timing = 0;
    /* the plan's own second stream is not forked inside the capture: that
     * nested fork ended the process in the HIP runtime on this stack */
    pl->capturing = 1;
    size_t o = (size_t)S->off[p] * S->N;
    rc = fmbh_stream_wait(S->bs[p], S->ev_fork);
    if (!rc) rc = plan_apply(pl, S->trans, prec, S->xs + o, S->ys + o, S->nr[p], S->bs[p]);
    pl->timing = tm;
    pl->capturing = 0;
    if (!rc) rc = fmbh_event_record(S->ev_join[p], S->bs[p]);
    if (!rc) rc = fmbh_stream_wait(S->cs, S->ev_join[p]);
  }
  rc2 = fmbh_capture_end(S->cs, exec);
  return rc ? rc : rc2;
}

/* y = A x (A^T x), every plan on its columns; y must be S->ys in graph mode */
static int apply(sys_t *S, int prec, const fmb_c64 *x, fmb_c64 *y, void *stream) {
  if (!S->graph) {
    for (int p = 0; p < S->np; p++) {
      if (!S->live[p]) continue;
      size_t o = (size_t)S->off[p] * S->N;
      int rc = plan_apply(S->pl[p], S->trans, prec, x + o, y + o, S->nr[p], stream);
      if (rc) return rc;
    }
    return FMB_OK;
  }
  int i = prec == FMB_MIXED ? 1 : 0, rc;
  if (!S->exec[i] && (rc = graph_capture(S, prec, &S->exec[i]))) return rc;
  if (x != S->xs &&
      (rc = fmbh_d2d(S->xs, x, sizeof(fmb_c64) * (size_t)S->off[S->np] * S->N, stream)))
    return rc;
  return fmbh_graph_launch(S->exec[i], stream);
}

/* validates the plans of a system and fills S (columns, live plans) */
static int sys_init(sys_t *S, fmb_plan *const *pls, const int *nrs, int np, int trans,
                    int graph) {
  memset(S, 0, sizeof(*S));
  if (!pls || !nrs || np < 1)
    return fmb_set_error(FMB_ERR_ARG, "fmb_gmres: invalid arguments");
  for (int p = 0; p < np; p++) {
    const fmb_plan *q = pls[p];
    if (!q || !q->ready || nrs[p] < 1 || nrs[p] > q->nrhs_max ||
        q->tree->nelem != pls[0]->tree->nelem || q->work_prec != pls[0]->work_prec)
      return fmb_set_error(FMB_ERR_ARG, "fmb_gmres: plan %d does not fit the batch", p);
    if (trans && q->prm.rcr_path == FMB_RCR_VALU)
      return fmb_set_error(FMB_ERR_ARG, "transposed solve needs rcr_path FMB_RCR_AUTO");
    /* the branches of a captured batch run concurrently, and a plan's pass
     * owns its staging vectors, coefficients and workspace: a plan listed
     * twice would race with itself */
    for (int r = 0; graph && r < p; r++)
      if (pls[r] == q)
        return fmb_set_error(FMB_ERR_ARG,
                             "fmb_gmres: plan %d is plan %d again (a graph batch needs "
                             "distinct plans)", p, r);
  }
  S->pl = pls;
  S->nr = nrs;
  S->np = np;
  S->trans = trans;
  S->graph = graph;
  S->N = pls[0]->tree->nelem;
  S->off = (int *)malloc(sizeof(int) * (np + 1));
  S->live = (int *)malloc(sizeof(int) * np);
  if (!S->off || !S->live) return fmb_set_error(FMB_ERR_NOMEM, "GMRES system");
  S->off[0] = 0;
  for (int p = 0; p < np; p++) {
    S->off[p + 1] = S->off[p] + nrs[p];
    S->live[p] = 1;
  }
  return FMB_OK;
}

/* graph-mode resources: staging input, the output vector ys, streams */
static int sys_graph_init(sys_t *S, fmb_c64 *ys) {
  int rc = fmbh_malloc((void **)&S->xs, sizeof(fmb_c64) * (size_t)S->off[S->np] * S->N);
  if (rc) return rc;
  S->ys = ys;
  S->bs = (void **)calloc(S->np, sizeof(void *));
  S->ev_join = (void **)calloc(S->np, sizeof(void *));
  if (!S->bs || !S->ev_join) return fmb_set_error(FMB_ERR_NOMEM, "GMRES graph branches");
  if ((rc = fmbh_stream_create(&S->cs))) return rc;
  if ((rc = fmbh_event_create(&S->ev_fork))) return rc;
  for (int p = 0; p < S->np; p++) {
    if ((rc = fmbh_stream_create(&S->bs[p]))) return rc;
    if ((rc = fmbh_event_create(&S->ev_join[p]))) return rc;
  }
  return FMB_OK;
}

static void sys_free(sys_t *S) {
  if (S->cs) fmbh_sync(S->cs);
  graphs_drop(S);
  for (int p = 0; p < S->np; p++) {
    if (S->bs) fmbh_stream_destroy(S->bs[p]);
    if (S->ev_join) fmbh_event_destroy(S->ev_join[p]);
  }
  fmbh_stream_destroy(S->cs);
  fmbh_event_destroy(S->ev_fork);
  fmbh_free(S->xs);
  free(S->bs);
  free(S->ev_join);
  free(S->off);
  free(S->live);
  memset(S, 0, sizeof(*S));
}

/* x += V y over the first j Arnoldi vectors, y from the triangular solve of
 * the rotated Hessenberg systems, for the columns with sel[r] (others y = 0) */
static int update_x(const zc *Hh, const zc *g, zc *yv, fmb_c64 *hh, int m, int nrhs,
                    int j, const int *sel, const fmb_c64 *dV, fmb_c64 *dy, fmb_c64 *x,
                    int N, void *stream) {
#define HQ(r, i, c) Hh[((size_t)(r) * m + (c)) * (m + 1) + (i)]
  for (int r = 0; r < nrhs; r++) {
    for (int q = j - 1; q >= 0; q--) {
      zc s = g[(size_t)r * (m + 1) + q];
      for (int c2 = q + 1; c2 < j; c2++) s -= HQ(r, q, c2) * yv[(size_t)r * (m + 1) + c2];
      zc d = HQ(r, q, q);
      yv[(size_t)r * (m + 1) + q] = (sel[r] && cabs(d) > 0) ? s / d : 0.0;
    }
    for (int q = 0; q < j; q++) hh[(size_t)q * nrhs + r] = fromz(yv[(size_t)r * (m + 1) + q]);
  }
#undef HQ
  int rc = fmbh_h2d(dy, hh, sizeof(fmb_c64) * j * nrhs, stream);
  if (!rc) rc = fmbh_zgemv_add(dV, j, dy, x, nrhs, N, stream);
  if (!rc) rc = fmbh_sync(stream); /* hh is reused at once */
  return rc;
}

/* est_stop: a cycle whose estimated residuals all meet the tolerance ends
 * the solve without the FP64 true residual (for a caller that checks the
 * result itself); relres then holds those estimates */
static int gmres_run(fmb_plan *const *pls, const int *nrs, int np, int trans, int graph,
                     const fmb_c64 *b, fmb_c64 *x, int restart, double tol, int maxit,
                     int est_stop, fmb_gmres_info *info, void *stream) {
  if (!b || !x || restart < 1 || !(tol > 0) || maxit < 1)
    return fmb_set_error(FMB_ERR_ARG, "fmb_gmres: invalid arguments");
  sys_t S;
  int rc = sys_init(&S, pls, nrs, np, trans, graph);
  if (rc) {
    sys_free(&S);
    return rc;
  }
  const fmb_plan *pl = pls[0];
  const int nrhs = S.off[np];
  double t0 = now_s();
  const int N = pl->tree->nelem, m = restart;
  const int work = pl->work_prec;
  const double itol = work == FMB_MIXED ? 0.5 * tol : tol; /* Arnoldi stop */
  size_t vbytes = sizeof(fmb_c64) * (size_t)N * nrhs;
  fmb_c64 *dV = NULL, *dw = NULL, *dr = NULL, *dh = NULL, *dh2 = NULL, *dy = NULL;
  double *dn = NULL;
  zc *Hh = NULL, *cs = NULL, *sn = NULL, *g = NULL, *yv = NULL;
  fmb_c64 *hh = NULL, *hh2 = NULL;
  double *bn = NULL, *rn = NULL, *hn = NULL, *rsave = NULL;
  int *active = NULL, *sel = NULL;
  int iters = 0, mv64 = 0, mvw = 0, restarts = 0, conv = 0;
  if ((rc = fmbh_malloc((void **)&dV, vbytes * (m + 1)))) goto done;
  if ((rc = fmbh_malloc((void **)&dw, vbytes))) goto done;
  if ((rc = fmbh_malloc((void **)&dr, vbytes))) goto done;
  if ((rc = fmbh_malloc((void **)&dh, sizeof(fmb_c64) * (m + 1) * nrhs))) goto done;
  if ((rc = fmbh_malloc((void **)&dh2, sizeof(fmb_c64) * (m + 1) * nrhs))) goto done;
  if ((rc = fmbh_malloc((void **)&dy, sizeof(fmb_c64) * (m + 1) * nrhs))) goto done;
  if ((rc = fmbh_malloc((void **)&dn, sizeof(double) * nrhs))) goto done;
  if (graph && (rc = sys_graph_init(&S, dw))) goto done;
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
  rsave = (double *)malloc(sizeof(double) * nrhs);
  active = (int *)malloc(sizeof(int) * nrhs);
  sel = (int *)malloc(sizeof(int) * nrhs);
  if (!Hh || !cs || !sn || !g || !yv || !hh || !hh2 || !bn || !rn || !hn || !rsave || !active ||
      !sel) {
    rc = fmb_set_error(FMB_ERR_NOMEM, "GMRES host workspace");
    goto done;
  }
#define HH(r, i, j) Hh[((size_t)(r) * m + (j)) * (m + 1) + (i)]
  if ((rc = fmbh_znorms(x, nrhs, N, dn, stream))) goto done;
  if ((rc = fmbh_d2h(bn, dn, sizeof(double) * nrhs, stream))) goto done;
  int x_zero = 1;
  for (int r = 0; r < nrhs; r++) {
    /* a NaN compares false with every tolerance and would read as
     * converged: non-finite inputs are refused */
    if (!isfinite(bn[r])) {
      rc = fmb_set_error(FMB_ERR_ARG, "fmb_gmres: initial guess %d holds NaN or Inf", r);
      goto done;
    }
    x_zero &= bn[r] == 0.0;
  }
  if ((rc = fmbh_znorms(b, nrhs, N, dn, stream))) goto done;
  if ((rc = fmbh_d2h(bn, dn, sizeof(double) * nrhs, stream))) goto done;
  for (int r = 0; r < nrhs; r++) {
    if (!isfinite(bn[r])) {
      rc = fmb_set_error(FMB_ERR_ARG, "fmb_gmres: right-hand side %d holds NaN or Inf", r);
      goto done;
    }
    if (bn[r] == 0.0) bn[r] = 1.0;
  }
  while (1) {
    /* true residual r = b - A x in FP64 (A x = 0 for a zero first guess) */
    if (x_zero) {
      if ((rc = fmbh_memset(dw, 0, vbytes, stream))) goto done;
      x_zero = 0;
    } else {
      if ((rc = apply(&S, FMB_FP64, x, dw, stream))) goto done;
      mv64++;
    }
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
    /* plans out of the graphs are not applied any more: their residuals are
     * the ones they converged with (their x no longer changes) */
    for (int p = 0; p < np; p++)
      if (!S.live[p])
        for (int r = S.off[p]; r < S.off[p + 1]; r++) rn[r] = rsave[r];
    conv = 1;
    for (int r = 0; r < nrhs; r++) {
      if (!isfinite(rn[r])) {
        rc = fmb_set_error(FMB_ERR_NOCONV,
                           "GMRES: true residual of right-hand side %d is not finite "
                           "after %d steps", r, iters);
        goto done;
      }
      active[r] = rn[r] / bn[r] > tol;
      if (active[r]) conv = 0;
    }
    if (conv || iters >= maxit) break;
    if (np > 1)
      for (int p = 0; p < np; p++) {
        int any = 0;
        for (int r = S.off[p]; r < S.off[p + 1]; r++) any |= active[r];
        if (S.live[p] && !any) {
          /* converged plan: out of the graphs, its columns of w held at 0 */
          S.live[p] = 0;
          for (int r = S.off[p]; r < S.off[p + 1]; r++) rsave[r] = rn[r];
          graphs_drop(&S);
          size_t o = (size_t)S.off[p] * S.N;
          if ((rc = fmbh_memset(dw + o, 0, sizeof(fmb_c64) * nrs[p] * (size_t)S.N, stream)))
            goto done;
        }
      }
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
    int j, cyc_done = 0;
    for (j = 0; j < m && iters < maxit; j++) {
      fmb_c64 *Vj = dV + (size_t)j * N * nrhs;
      if ((rc = apply(&S, work, Vj, dw, stream))) goto done;
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
        if (active[r] && cabs(g[(size_t)r * (m + 1) + j + 1]) / bn[r] > itol)
          all_done = 0;
      }
      if ((rc = fmbh_h2d(dh, hh, sizeof(fmb_c64) * nrhs, stream))) goto done;
      if ((rc = fmbh_zaxpby(dh, dw, NULL, dV + (size_t)(j + 1) * N * nrhs, nrhs,
                            N, stream)))
        goto done;
      /* a batch retires a plan as soon as its estimated residuals meet the
       * tolerance: its columns of x take this step's update, its FP64 true
       * residual is checked, and it leaves the pass (a plan that fails the
       * check rejoins at the next restart) */
      for (int p = 0; p < np && np > 1 && !all_done; p++) {
        int any = 0, ok = 1;
        for (int r = S.off[p]; r < S.off[p + 1]; r++)
          if (active[r]) {
            any = 1;
            if (cabs(g[(size_t)r * (m + 1) + j + 1]) / bn[r] > itol) ok = 0;
          }
        if (!S.live[p] || !any || !ok) continue;
        for (int r = 0; r < nrhs; r++)
          sel[r] = r >= S.off[p] && r < S.off[p + 1] && active[r];
        if ((rc = update_x(Hh, g, yv, hh, m, nrhs, j + 1, sel, dV, dy, x, N, stream)))
          goto done;
        for (int r = S.off[p]; r < S.off[p + 1]; r++) active[r] = 0;
        const size_t o = (size_t)S.off[p] * N;
        const int nq = nrs[p];
        if (est_stop) {
          /* the caller checks the result: retire on the estimates */
          for (int r = S.off[p]; r < S.off[p + 1]; r++)
            rsave[r] = sel[r] ? cabs(g[(size_t)r * (m + 1) + j + 1]) : rn[r];
          S.live[p] = 0;
          graphs_drop(&S);
          if ((rc = fmbh_memset(dw + o, 0, sizeof(fmb_c64) * nq * (size_t)N, stream)))
            goto done;
          continue;
        }
        /* w's columns of p are free (V_{j+1} is formed): A x_p there */
        if ((rc = plan_apply(S.pl[p], trans, FMB_FP64, x + o, dw + o, nq, stream))) goto done;
        mv64++;
        if ((rc = fmbh_d2d(dr + o, b + o, sizeof(fmb_c64) * nq * (size_t)N, stream))) goto done;
        for (int r = 0; r < nq; r++) {
          hh[r].re = 1.0;
          hh[r].im = 0.0;
          hh2[r].re = -1.0;
          hh2[r].im = 0.0;
        }
        if ((rc = fmbh_h2d(dh, hh, sizeof(fmb_c64) * nq, stream))) goto done;
        if ((rc = fmbh_h2d(dh2, hh2, sizeof(fmb_c64) * nq, stream))) goto done;
        /* r_p = 1 * b_p + (-1) * A x_p */
        if ((rc = fmbh_zaxpby(dh2, dw + o, dh, dr + o, nq, N, stream))) goto done;
        if ((rc = fmbh_znorms(dr + o, nq, N, dn, stream))) goto done;
        if ((rc = fmbh_d2h(hn, dn, sizeof(double) * nq, stream))) goto done;
        int pass = 1;
        for (int r = 0; r < nq; r++) pass &= hn[r] / bn[S.off[p] + r] <= tol;
        if (!pass) continue;
        S.live[p] = 0;
        for (int r = 0; r < nq; r++) rsave[S.off[p] + r] = hn[r];
        graphs_drop(&S);
        if ((rc = fmbh_memset(dw + o, 0, sizeof(fmb_c64) * nq * (size_t)N, stream))) goto done;
      }
      if (all_done) {
        j++;
        cyc_done = 1;
        break;
      }
    }
    /* x += V y, y from the triangular solve (inactive rhs get y = 0) */
    if ((rc = update_x(Hh, g, yv, hh, m, nrhs, j, active, dV, dy, x, N, stream))) goto done;
    if (est_stop && cyc_done) {
      for (int r = 0; r < nrhs; r++)
        if (active[r]) rn[r] = cabs(g[(size_t)r * (m + 1) + j]);
      for (int p = 0; p < np; p++)
        if (!S.live[p])
          for (int r = S.off[p]; r < S.off[p + 1]; r++) rn[r] = rsave[r];
      conv = 1;
      break;
    }
  }
  rc = conv ? FMB_OK : fmb_set_error(FMB_ERR_NOCONV,
                                     "GMRES: tolerance not reached in %d steps",
                                     iters);
  if (info)
    for (int p = 0; p < np; p++) {
      fmb_gmres_info *ip = info + p;
      memset(ip, 0, sizeof(*ip));
      ip->iters = iters;
      ip->matvecs_fp64 = mv64;
      ip->matvecs_work = mvw;
      ip->restarts = restarts;
      for (int r = 0; r < nrs[p] && r < 64; r++)
        ip->relres[r] = rn[S.off[p] + r] / bn[S.off[p] + r];
    }
#undef HH
done:
  if (info)
    for (int p = 0; p < np; p++) info[p].seconds = now_s() - t0;
  fmbh_sync(stream);
  sys_free(&S);
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
  free(rsave);
  free(active);
  free(sel);
  return rc;
}

int fmb_gmres(fmb_plan *pl, const fmb_c64 *b, fmb_c64 *x, int nrhs,
              int restart, double tol, int maxit, fmb_gmres_info *info,
              void *stream) {
  return gmres_run(&pl, &nrhs, 1, 0, 0, b, x, restart, tol, maxit, 0, info, stream);
}

int fmb_gmres_transpose(fmb_plan *pl, const fmb_c64 *b, fmb_c64 *x, int nrhs,
                        int restart, double tol, int maxit, fmb_gmres_info *info,
                        void *stream) {
  return gmres_run(&pl, &nrhs, 1, 1, 0, b, x, restart, tol, maxit, 0, info, stream);
}

int fmb_gmres_batch(fmb_plan *const *plans, const int *nrhs, int nplans, int graph,
                    const fmb_c64 *b, fmb_c64 *x, int restart, double tol, int maxit,
                    fmb_gmres_info *info, void *stream) {
  return gmres_run(plans, nrhs, nplans, 0, graph, b, x, restart, tol, maxit, 0, info, stream);
}

/* Robin coefficients of two plans on one tree agree (both rigid or equal) */
static int same_robin(const fmb_plan *a, const fmb_plan *b, int N, void *stream, int *same) {
  *same = 0;
  if (!a->d_robin || !b->d_robin) {
    *same = !a->d_robin && !b->d_robin;
    return FMB_OK;
  }
  fmb_c64 *ha = (fmb_c64 *)malloc(sizeof(fmb_c64) * 2 * (size_t)N);
  if (!ha) return fmb_set_error(FMB_ERR_NOMEM, "Robin coefficient copies");
  int rc = fmbh_d2h(ha, a->d_robin, sizeof(fmb_c64) * (size_t)N, stream);
  if (!rc) rc = fmbh_d2h(ha + N, b->d_robin, sizeof(fmb_c64) * (size_t)N, stream);
  if (!rc) *same = !memcmp(ha, ha + N, sizeof(fmb_c64) * (size_t)N);
  free(ha);
  return rc;
}

/* the inner plan of a refined solve agrees with its outer plan */
static int refined_pair_check(const fmb_plan *outer, const fmb_plan *inner, int nr, void *stream) {
  if (!outer || !inner || nr < 1)
    return fmb_set_error(FMB_ERR_ARG, "fmb_gmres_refined: invalid arguments");
  if (nr > outer->nrhs_max || nr > inner->nrhs_max)
    return fmb_set_error(FMB_ERR_RANGE,
                         "fmb_gmres_refined: %d right-hand sides above a plan's nrhs_max", nr);
  if (inner->tree != outer->tree || fabs(inner->k - outer->k) > 1e-12 * fabs(outer->k) ||
      inner->alpha_re != outer->alpha_re || inner->alpha_im != outer->alpha_im)
    return fmb_set_error(FMB_ERR_ARG,
                         "fmb_gmres_refined: the inner plan must share the outer plan's tree, "
                         "wavenumber and coupling");
  int same = 0, rc = same_robin(outer, inner, outer->tree->nelem, stream, &same);
  if (rc) return rc;
  if (!same)
    return fmb_set_error(FMB_ERR_ARG, "fmb_gmres_refined: the plans' Robin coefficients differ");
  return FMB_OK;
}

/* Iterative refinement on each outer plan's FP64 residual around GMRES on
 * its inner plan, a plan of the same tree, wavenumber, coupling and Robin
 * coefficients (typically built with a looser truncation, so its passes are
 * cheaper).  Plan p owns columns [off[p], off[p+1]) of b and x.  Each cycle
 * solves A_in d = r to inner_tol relative to |r| (GMRES on the inner plans in
 * lock step, in their work precision, stopped on the estimated residual: the
 * outer FP64 residual checks every cycle) and sets x += d.  Right-hand sides
 * whose outer residual meets tol sit out with r = 0, and a plan whose
 * right-hand sides all do needs no further outer pass.  The contraction per
 * cycle is about inner_tol plus the inner operator's relative difference
 * from the outer one. */
static int refined_run(fmb_plan *const *outers, fmb_plan *const *inners, const int *nrs, int np,
                       int trans, int graph, const fmb_c64 *b, fmb_c64 *x, int restart,
                       double tol, double inner_tol, int maxit, fmb_gmres_info *info,
                       void *stream) {
  if (!outers || !inners || !nrs || np < 1 || !b || !x || restart < 1 || !(tol > 0) ||
      !(inner_tol > 0 && inner_tol < 1) || maxit < 1)
    return fmb_set_error(FMB_ERR_ARG, "fmb_gmres_refined: invalid arguments");
  int rc, nrhs = 0;
  for (int p = 0; p < np; p++) {
    if ((rc = refined_pair_check(outers[p], inners[p], nrs[p], stream))) return rc;
    if (outers[p]->tree->nelem != outers[0]->tree->nelem)
      return fmb_set_error(FMB_ERR_ARG, "fmb_gmres_refined: plans of different element counts");
    nrhs += nrs[p];
  }
  const int N = outers[0]->tree->nelem;
  double t0 = now_s();
  const size_t vbytes = sizeof(fmb_c64) * (size_t)N * nrhs;
  fmb_c64 *dw = NULL, *dr = NULL, *dd = NULL, *dh = NULL, *dh2 = NULL;
  double *dn = NULL, *bn = NULL, *rn = NULL, *tn = NULL;
  fmb_c64 *one = NULL, *mone = NULL;
  fmb_gmres_info *ii = NULL;
  int *off = NULL, *pdone = NULL, *mvin = NULL, *mv64 = NULL;
  int iters = 0, cycles = 0, conv = 0, x_zero = 1;
  if ((rc = fmbh_malloc((void **)&dw, vbytes))) goto done;
  if ((rc = fmbh_malloc((void **)&dr, vbytes))) goto done;
  if ((rc = fmbh_malloc((void **)&dd, vbytes))) goto done;
  if ((rc = fmbh_malloc((void **)&dh, sizeof(fmb_c64) * nrhs))) goto done;
  if ((rc = fmbh_malloc((void **)&dh2, sizeof(fmb_c64) * nrhs))) goto done;
  if ((rc = fmbh_malloc((void **)&dn, sizeof(double) * nrhs))) goto done;
  bn = (double *)malloc(sizeof(double) * nrhs);
  rn = (double *)malloc(sizeof(double) * nrhs);
  tn = (double *)malloc(sizeof(double) * nrhs);
  one = (fmb_c64 *)malloc(sizeof(fmb_c64) * nrhs);
  mone = (fmb_c64 *)malloc(sizeof(fmb_c64) * nrhs);
  ii = (fmb_gmres_info *)calloc((size_t)np, sizeof(fmb_gmres_info));
  off = (int *)malloc(sizeof(int) * (np + 1));
  pdone = (int *)calloc((size_t)np, sizeof(int));
  mvin = (int *)calloc((size_t)np, sizeof(int));
  mv64 = (int *)calloc((size_t)np, sizeof(int));
  if (!bn || !rn || !tn || !one || !mone || !ii || !off || !pdone || !mvin || !mv64) {
    rc = fmb_set_error(FMB_ERR_NOMEM, "fmb_gmres_refined host workspace");
    goto done;
  }
  off[0] = 0;
  for (int p = 0; p < np; p++) off[p + 1] = off[p] + nrs[p];
  for (int r = 0; r < nrhs; r++) {
    one[r].re = 1.0;
    one[r].im = 0.0;
    mone[r].re = -1.0;
    mone[r].im = 0.0;
  }
  if ((rc = fmbh_h2d(dh, one, sizeof(fmb_c64) * nrhs, stream))) goto done;
  if ((rc = fmbh_h2d(dh2, mone, sizeof(fmb_c64) * nrhs, stream))) goto done;
  if ((rc = fmbh_znorms(x, nrhs, N, dn, stream))) goto done;
  if ((rc = fmbh_d2h(rn, dn, sizeof(double) * nrhs, stream))) goto done;
  for (int r = 0; r < nrhs; r++) {
    if (!isfinite(rn[r])) {
      rc = fmb_set_error(FMB_ERR_ARG, "fmb_gmres_refined: initial guess %d holds NaN or Inf", r);
      goto done;
    }
    x_zero &= rn[r] == 0.0;
  }
  if ((rc = fmbh_znorms(b, nrhs, N, dn, stream))) goto done;
  if ((rc = fmbh_d2h(bn, dn, sizeof(double) * nrhs, stream))) goto done;
  for (int r = 0; r < nrhs; r++) {
    if (!isfinite(bn[r])) {
      rc = fmb_set_error(FMB_ERR_ARG, "fmb_gmres_refined: right-hand side %d holds NaN or Inf",
                         r);
      goto done;
    }
    if (bn[r] == 0.0) bn[r] = 1.0;
  }
  while (1) {
    /* r = b - A_out x in FP64 (r = b for a zero first guess); converged
     * plans keep r = 0 */
    if ((rc = fmbh_d2d(dr, b, vbytes, stream))) goto done;
    if (!x_zero) {
      for (int p = 0; p < np; p++) {
        if (pdone[p]) continue;
        const size_t o = (size_t)off[p] * N;
        if ((rc = plan_apply(outers[p], trans, FMB_FP64, x + o, dw + o, nrs[p], stream)))
          goto done;
        mv64[p]++;
        if ((rc = fmbh_zaxpby(dh2 + off[p], dw + o, dh + off[p], dr + o, nrs[p], N, stream)))
          goto done;
      }
    }
    if ((rc = fmbh_znorms(dr, nrhs, N, dn, stream))) goto done;
    if ((rc = fmbh_d2h(tn, dn, sizeof(double) * nrhs, stream))) goto done;
    /* finished plans keep the norms they converged with */
    for (int p = 0; p < np; p++)
      if (!pdone[p])
        for (int r = off[p]; r < off[p + 1]; r++) rn[r] = tn[r];
    conv = 1;
    for (int r = 0; r < nrhs; r++)
      if (!isfinite(rn[r])) {
        rc = fmb_set_error(FMB_ERR_NOCONV,
                           "fmb_gmres_refined: residual of right-hand side %d is not finite "
                           "after %d inner steps", r, iters);
        goto done;
      }
    for (int p = 0; p < np; p++) {
      int ok = 1;
      for (int r = off[p]; r < off[p + 1]; r++) ok &= rn[r] / bn[r] <= tol;
      pdone[p] |= ok;
      conv &= ok;
    }
    if (conv || iters >= maxit) break;
    for (int r = 0; r < nrhs; r++)
      if (rn[r] / bn[r] <= tol &&
          (rc = fmbh_memset(dr + (size_t)r * N, 0, sizeof(fmb_c64) * (size_t)N, stream)))
        goto done;
    if ((rc = fmbh_memset(dd, 0, vbytes, stream))) goto done;
    memset(ii, 0, sizeof(fmb_gmres_info) * np);
    rc = gmres_run(inners, nrs, np, trans, graph, dr, dd, restart, inner_tol, maxit - iters, 1,
                   ii, stream);
    if (rc && rc != FMB_ERR_NOCONV) goto done;
    iters += ii[0].iters;
    for (int p = 0; p < np; p++) mvin[p] += ii[p].matvecs_fp64 + ii[p].matvecs_work;
    cycles++;
    /* x = 1 * d + 1 * x */
    if ((rc = fmbh_zaxpby(dh, dd, dh, x, nrhs, N, stream))) goto done;
    x_zero = 0;
  }
  rc = conv ? FMB_OK
            : fmb_set_error(FMB_ERR_NOCONV,
                            "fmb_gmres_refined: tolerance not reached in %d inner steps", iters);
  if (info)
    for (int p = 0; p < np; p++) {
      fmb_gmres_info *ip = info + p;
      memset(ip, 0, sizeof(*ip));
      ip->iters = iters;
      ip->matvecs_fp64 = mv64[p];
      ip->matvecs_work = mvin[p];
      ip->restarts = cycles;
      for (int r = 0; r < nrs[p] && r < 64; r++) ip->relres[r] = rn[off[p] + r] / bn[off[p] + r];
    }
done:
  if (info)
    for (int p = 0; p < np; p++) info[p].seconds = now_s() - t0;
  fmbh_sync(stream);
  fmbh_free(dw);
  fmbh_free(dr);
  fmbh_free(dd);
  fmbh_free(dh);
  fmbh_free(dh2);
  fmbh_free(dn);
  free(bn);
  free(rn);
  free(tn);
  free(one);
  free(mone);
  free(ii);
  free(off);
  free(pdone);
  free(mvin);
  free(mv64);
  return rc;
}

int fmb_gmres_refined(fmb_plan *outer, fmb_plan *inner, const fmb_c64 *b, fmb_c64 *x,
                      int nrhs, int restart, double tol, double inner_tol, int maxit,
                      fmb_gmres_info *info, void *stream) {
  return refined_run(&outer, &inner, &nrhs, 1, 0, 0, b, x, restart, tol, inner_tol, maxit, info,
                     stream);
}

int fmb_gmres_refined_transpose(fmb_plan *outer, fmb_plan *inner, const fmb_c64 *b,
                                fmb_c64 *x, int nrhs, int restart, double tol,
                                double inner_tol, int maxit, fmb_gmres_info *info,
                                void *stream) {
  return refined_run(&outer, &inner, &nrhs, 1, 1, 0, b, x, restart, tol, inner_tol, maxit, info,
                     stream);
}

int fmb_gmres_batch_refined(fmb_plan *const *outers, fmb_plan *const *inners, const int *nrhs,
                            int nplans, int graph, const fmb_c64 *b, fmb_c64 *x, int restart,
                            double tol, double inner_tol, int maxit, fmb_gmres_info *info,
                            void *stream) {
  return refined_run(outers, inners, nrhs, nplans, 0, graph, b, x, restart, tol, inner_tol, maxit,
                     info, stream);
}

int fmb_matvec_batch(fmb_plan *const *plans, const int *nrhs, int nplans, int graph,
                     int prec, const fmb_c64 *x, fmb_c64 *y, int reps, void *stream) {
  if (!x || !y || reps < 1)
    return fmb_set_error(FMB_ERR_ARG, "fmb_matvec_batch: invalid arguments");
  sys_t S;
  int rc = sys_init(&S, plans, nrhs, nplans, 0, graph);
  fmb_c64 *ys = NULL;
  const size_t nv = (size_t)(rc ? 0 : S.off[S.np]) * S.N;
  if (!rc && graph) {
    if (!(rc = fmbh_malloc((void **)&ys, sizeof(fmb_c64) * nv))) rc = sys_graph_init(&S, ys);
  }
  for (int r = 0; r < reps && !rc; r++) {
    fmb_c64 *out = y + (size_t)r * nv;
    rc = apply(&S, prec, x, graph ? ys : out, stream);
    if (!rc && graph) rc = fmbh_d2d(out, ys, sizeof(fmb_c64) * nv, stream);
    if (!rc) rc = fmbh_sync(stream);
  }
  fmbh_sync(stream);
  sys_free(&S);
  fmbh_free(ys);
  return rc;
}
