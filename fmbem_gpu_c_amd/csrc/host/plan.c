/* plan.c -- per-frequency plan (the operator half of genOctree,
 * translation.cu:3078-3348, plus bases and near field) and the FMM pass
 * (OprL, translation.cu:3412-3769, completed with L2P and near field).
 *
 * A plan has three tiers of state:
 *   geometry  -- translation directions and their polar-angle classes, octant
 *                geometry, M2L work chunks, quadrature rules, near-field
 *                layout: built once per tree;
 *   rotation  -- real rotation blocks H_n(beta) of every polar angle at order
 *                Pmax, padded copies and azimuth phase tables: independent of
 *                the wavenumber, rebuilt only when an order grows past Pmax
 *                (genOctree builds rotMat1/rotMat2 once for the largest p,
 *                translation.cu:3160-3174);
 *   frequency -- orders, scale factors, coaxial matrices, bases, near-field
 *                blocks, Robin coefficients and translation schedules:
 *                rebuilt by fmb_plan_retune (genOctree regenerates only the
 *                coaxial matrices per wavenumber, translation.cu:3176-3186,
 *                3294-3302). */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../fmb_internal.h"
#include "../hip/fmb_hip.h"
#include "host.h"

void fmb_plan_params_default(fmb_plan_params *p) {
  memset(p, 0, sizeof(*p));
  p->precision = FMB_FP64;
  p->p_rule = FMB_PRULE_TRUNCNUM;
  p->eps = 0.05;   /* initOctree eps (translation.cu:2991) */
  p->sigma = 1.5;  /* truncNum sigma at every call (translation.cu:3432) */
  p->p_min = 1;
  p->quad_rule = FMB_RULE_GAUSS7;
  p->quad_order = 3;
  p->near_rule = FMB_RULE_COLLAPSED;
  p->near_order = 8;
  p->near_ratio = 2.0;
  p->nedge = 24;
  p->nrhs_max = 1;
  p->m2l_tile = 256; /* measured: level-7 M2L 70.0 -> 66.5 ms (FP32), 128 -> 116 ms (FP64) */
  p->m2l_tile_min_mb = 256;
}

static const int k_dist2[FMB_NDIST] = {4, 5, 6, 8, 9, 10, 11, 12,
                                       13, 14, 17, 18, 19, 22, 27};

static int dist_class(int d2) {
  for (int i = 0; i < FMB_NDIST; i++)
    if (k_dist2[i] == d2) return i;
  return -1;
}

static int beta_of(fmb_plan *pl, double cb) {
  for (int i = 0; i < pl->nbeta; i++)
    if (fabs(pl->beta_cos[i] - cb) < 1e-12) return i;
  pl->beta_cos[pl->nbeta] = cb;
  return pl->nbeta++;
}

static int order_for(const fmb_plan_params *prm, double k, double side, int l) {
  int p;
  if (prm->p_rule == FMB_PRULE_EXPLICIT)
    p = prm->p_levels[l];
  else if (prm->p_rule == FMB_PRULE_EXCESS) {
    double ka = k * sqrt(3.0) * side;
    p = (int)ceil(ka + 1.8 * log10(ka + M_PI));
  } else
    p = fmb_trunc_num(k, prm->eps, prm->sigma, side);
  if (p < prm->p_min) p = prm->p_min;
  return p;
}

/* orders of levels 2..lmax at wavenumber k and their maximum */
static int level_orders(const fmb_plan *pl, double k, int *p, int *pmax) {
  const fmb_tree *t = pl->tree;
  *pmax = 0;
  for (int l = 2; l <= t->lmax; l++) {
    p[l] = order_for(&pl->prm, k, t->D * ldexp(1.0, -l), l);
    if (p[l] < 1 || p[l] > 200)
      return fmb_set_error(FMB_ERR_RANGE, "order %d at level %d (k = %g)", p[l], l, k);
    if (p[l] > *pmax) *pmax = p[l];
  }
  return FMB_OK;
}

/* wavenumber-dependent arrays of one opset */
static int opset_alloc(fmb_plan *pl, fmb_opset *o) {
  const fmb_tree *t = pl->tree;
  int rc, L = t->lmax, N = t->nelem, nr = pl->nrhs_max;
  for (int l = 2; l <= L; l++) {
    size_t cs = (size_t)fmb_coax_size(pl->p[l]) * FMB_NDIST * o->csz;
    if ((rc = fmbh_malloc(&o->d_Csr[l], cs))) return rc;
    pl->bytes += cs;
    if (l > 2) {
      int P = pl->p[l] > pl->p[l - 1] ? pl->p[l] : pl->p[l - 1];
      size_t c2 = (size_t)fmb_coax_size(P) * o->csz;
      if ((rc = fmbh_malloc(&o->d_Cm2m[l], c2))) return rc;
      if ((rc = fmbh_malloc(&o->d_Cl2l[l], c2))) return rc;
      pl->bytes += 2 * c2;
    }
    size_t ex = (size_t)t->nbox[l] * nr * pl->p[l] * pl->p[l] * o->csz;
    if ((rc = fmbh_malloc(&o->d_M[l], ex))) return rc;
    if ((rc = fmbh_malloc(&o->d_L[l], ex))) return rc;
    pl->bytes += 2 * ex;
  }
  size_t bz = (size_t)N * pl->p[L] * pl->p[L] * o->csz;
  if ((rc = fmbh_malloc(&o->d_P2M, bz))) return rc;
  if ((rc = fmbh_malloc(&o->d_L2P, bz))) return rc;
  pl->bytes += 2 * bz;
  size_t nz = pl->near_total * o->csz;
  if ((rc = fmbh_malloc(&o->d_near, nz))) return rc;
  pl->bytes += nz;
  size_t vz = (size_t)N * nr * o->csz;
  if ((rc = fmbh_malloc(&o->d_xs, vz))) return rc;
  if ((rc = fmbh_malloc(&o->d_ys, vz))) return rc;
  pl->bytes += 2 * vz;
  return FMB_OK;
}

static void opset_free_freq(fmb_opset *o) {
  for (int l = 0; l < FMB_MAXLEV; l++) {
    fmbh_free(o->d_Csr[l]);
    fmbh_free(o->d_Cm2m[l]);
    fmbh_free(o->d_Cl2l[l]);
    fmbh_free(o->d_M[l]);
    fmbh_free(o->d_L[l]);
    o->d_Csr[l] = o->d_Cm2m[l] = o->d_Cl2l[l] = o->d_M[l] = o->d_L[l] = NULL;
  }
  fmbh_free(o->d_P2M);
  fmbh_free(o->d_L2P);
  fmbh_free(o->d_near);
  fmbh_free(o->d_xs);
  fmbh_free(o->d_ys);
  fmbh_free(o->d_gws);
  o->d_P2M = o->d_L2P = o->d_near = o->d_xs = o->d_ys = o->d_gws = NULL;
  o->gws_grid = 0;
  o->gws_bytes = 0;
  o->gws_slot = 0;
}

static void opset_free_rot(fmb_opset *o) {
  fmbh_free(o->d_H);
  fmbh_free(o->d_Hpad);
  fmbh_free(o->d_Hb);
  o->d_H = o->d_Hpad = o->d_Hb = NULL;
  o->hb_half = 0;
  for (int i = 0; i < 3; i++) {
    fmbh_free(o->d_phase[i]);
    fmbh_free(o->d_phase_c[i]);
    o->d_phase[i] = o->d_phase_c[i] = NULL;
  }
  o->hpad_stride = 0;
}

static void plan_freq_free(fmb_plan *pl) {
  pl->ready = 0;
  fmbx_rcr_free(pl);
  for (int i = 0; i < 2; i++) opset_free_freq(&pl->ops[i]);
  for (int l = 0; l < FMB_MAXLEV; l++) {
    free(pl->h_scale[l]);
    fmbh_free(pl->d_scale[l]);
    fmbh_free(pl->d_tab_mmajor[l]);
    pl->h_scale[l] = NULL;
    pl->d_scale[l] = NULL;
    pl->d_tab_mmajor[l] = NULL;
  }
  fmbh_free(pl->d_robin);
  pl->d_robin = NULL;
  pl->bytes -= pl->bytes_freq;
  pl->bytes_freq = 0;
}

static void plan_rot_free(fmb_plan *pl) {
  for (int i = 0; i < 2; i++) opset_free_rot(&pl->ops[i]);
  pl->bytes -= pl->bytes_rot;
  pl->bytes_rot = 0;
  pl->Pmax = 0;
}

void fmb_plan_destroy(fmb_plan *pl) {
  if (!pl) return;
  plan_freq_free(pl);
  plan_rot_free(pl);
  for (int l = 0; l < FMB_MAXLEV; l++) fmbh_free(pl->d_m2l_chunks[l]);
  fmbh_free(pl->d_m2m_beta);
  fmbh_free(pl->d_l2l_beta);
  fmbh_free(pl->d_m2m_phi);
  fmbh_free(pl->d_l2l_phi);
  fmbh_free(pl->d_off_info);
  fmbh_free(pl->d_quad);
  fmbh_free(pl->d_quadn);
  fmbh_free(pl->d_edge);
  fmbh_free(pl->d_near_off);
  fmbh_free(pl->d_near_ncol);
  fmbh_free(pl->d_work64);
  if (pl->aux) fmbh_sync(pl->aux);
  fmbh_stream_destroy(pl->aux);
  fmbh_event_destroy(pl->ev_up);
  fmbh_event_destroy(pl->ev_aux);
  for (int l = 0; l < FMB_MAXLEV; l++) fmbh_event_destroy(pl->ev_m2l[l]);
  free(pl);
}

typedef struct {
  int beta, dist;
  double phi;
} off_info_t;

/* translation directions, octant geometry, M2L work chunks, quadrature rules
 * and near-field layout: everything fixed by the tree and the rules */
static int plan_geometry(fmb_plan *pl, void *stream) {
  const fmb_tree *t = pl->tree;
  const fmb_plan_params *prm = &pl->prm;
  int rc, L = t->lmax;
  /* directions: M2L offsets and the child/parent diagonals */
  pl->nbeta = 0;
  for (int o = 0; o < FMB_NOFF; o++) {
    int dx = o / 49 - 3, dy = (o / 7) % 7 - 3, dz = o % 7 - 3;
    int d2 = dx * dx + dy * dy + dz * dz;
    pl->off_beta[o] = pl->off_dist[o] = -1;
    pl->off_phi[o] = 0.0;
    if (dist_class(d2) < 0 || (abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1))
      continue;
    pl->off_beta[o] = beta_of(pl, dz / sqrt((double)d2));
    pl->off_dist[o] = dist_class(d2);
    pl->off_phi[o] = atan2((double)dy, (double)dx);
  }
  for (int oct = 0; oct < 8; oct++) {
    double d[3] = {((oct >> 2) & 1) - 0.5, ((oct >> 1) & 1) - 0.5, (oct & 1) - 0.5};
    double r = sqrt(0.75);
    pl->l2l_beta[oct] = beta_of(pl, d[2] / r);
    pl->l2l_phi[oct] = atan2(d[1], d[0]);
    pl->m2m_beta[oct] = beta_of(pl, -d[2] / r);
    pl->m2m_phi[oct] = atan2(-d[1], -d[0]);
  }
  {
    off_info_t info[FMB_NOFF];
    for (int o = 0; o < FMB_NOFF; o++) {
      info[o].beta = pl->off_beta[o] < 0 ? 0 : pl->off_beta[o];
      info[o].dist = pl->off_dist[o] < 0 ? 0 : pl->off_dist[o];
      info[o].phi = pl->off_phi[o];
    }
    if ((rc = fmbh_malloc(&pl->d_off_info, sizeof(info)))) return rc;
    if ((rc = fmbh_h2d(pl->d_off_info, info, sizeof(info), stream))) return rc;
    if ((rc = fmbh_malloc((void **)&pl->d_m2m_beta, sizeof(int) * 8))) return rc;
    if ((rc = fmbh_malloc((void **)&pl->d_l2l_beta, sizeof(int) * 8))) return rc;
    if ((rc = fmbh_malloc((void **)&pl->d_m2m_phi, sizeof(double) * 8))) return rc;
    if ((rc = fmbh_malloc((void **)&pl->d_l2l_phi, sizeof(double) * 8))) return rc;
    if ((rc = fmbh_h2d(pl->d_m2m_beta, pl->m2m_beta, sizeof(int) * 8, stream))) return rc;
    if ((rc = fmbh_h2d(pl->d_l2l_beta, pl->l2l_beta, sizeof(int) * 8, stream))) return rc;
    if ((rc = fmbh_h2d(pl->d_m2m_phi, pl->m2m_phi, sizeof(double) * 8, stream))) return rc;
    if ((rc = fmbh_h2d(pl->d_l2l_phi, pl->l2l_phi, sizeof(double) * 8, stream))) return rc;
    /* the h2d copies read the stack arrays asynchronously */
    if ((rc = fmbh_sync(stream))) return rc;
  }
  /* M2L work chunks per level: pairs of one offset bucket */
  for (int l = 2; l <= L; l++) {
    int np = t->npairs[l];
    int ch = np / 4096;
    if (ch < 1) ch = 1;
    if (ch > 16) ch = 16;
    int cap = np / ch + FMB_NOFF + 1, n = 0;
    int *c = (int *)malloc(sizeof(int) * 4 * cap);
    if (!c) return fmb_set_error(FMB_ERR_NOMEM, "M2L chunks");
    for (int o = 0; o < FMB_NOFF && np > 0; o++)
      for (int s = t->h_off_start[l][o]; s < t->h_off_start[l][o + 1]; s += ch) {
        int e = s + ch < t->h_off_start[l][o + 1] ? s + ch : t->h_off_start[l][o + 1];
        c[4 * n] = o;
        c[4 * n + 1] = s;
        c[4 * n + 2] = e;
        c[4 * n + 3] = 0;
        n++;
      }
    pl->n_m2l_chunks[l] = n;
    rc = fmbh_malloc(&pl->d_m2l_chunks[l], sizeof(int) * 4 * (n > 0 ? n : 1));
    if (!rc && n) rc = fmbh_h2d(pl->d_m2l_chunks[l], c, sizeof(int) * 4 * n, stream);
    if (!rc) rc = fmbh_sync(stream);
    free(c);
    if (rc) return rc;
  }
  /* quadrature tables */
  double *q = NULL;
  int nq = fmbx_tri_rule(prm->quad_rule, prm->quad_order, &q);
  if (nq < 1)
    return fmb_set_error(FMB_ERR_ARG, "panel rule %d order %d", prm->quad_rule,
                         prm->quad_order);
  pl->nq = nq;
  rc = fmbh_malloc((void **)&pl->d_quad, sizeof(double) * 4 * nq);
  if (!rc) rc = fmbh_h2d(pl->d_quad, q, sizeof(double) * 4 * nq, stream);
  if (!rc) rc = fmbh_sync(stream);
  free(q);
  if (rc) return rc;
  int nqn = fmbx_tri_rule(prm->near_rule, prm->near_order, &q);
  if (nqn < 1)
    return fmb_set_error(FMB_ERR_ARG, "close-pair rule %d order %d",
                         prm->near_rule, prm->near_order);
  pl->nqn = nqn;
  rc = fmbh_malloc((void **)&pl->d_quadn, sizeof(double) * 4 * nqn);
  if (!rc) rc = fmbh_h2d(pl->d_quadn, q, sizeof(double) * 4 * nqn, stream);
  if (!rc) rc = fmbh_sync(stream);
  free(q);
  if (rc) return rc;
  int ne = prm->nedge;
  double *gx = (double *)malloc(sizeof(double) * ne),
         *gw = (double *)malloc(sizeof(double) * ne),
         *e = (double *)malloc(sizeof(double) * 2 * ne);
  if (!gx || !gw || !e) {
    free(gx);
    free(gw);
    free(e);
    return fmb_set_error(FMB_ERR_NOMEM, "edge rule");
  }
  fmbx_gauss_legendre(ne, gx, gw);
  for (int i = 0; i < ne; i++) {
    e[2 * i] = gx[i];
    e[2 * i + 1] = gw[i];
  }
  rc = fmbh_malloc((void **)&pl->d_edge, sizeof(double) * 2 * ne);
  if (!rc) rc = fmbh_h2d(pl->d_edge, e, sizeof(double) * 2 * ne, stream);
  if (!rc) rc = fmbh_sync(stream);
  free(gx);
  free(gw);
  free(e);
  if (rc) return rc;
  /* the matrix-core L2L adds with atomics (the VALU one does not, so it
   * cannot share L_l with that level's M2L) */
  if (!prm->serial_levels && L >= 5 && prm->rcr_path != FMB_RCR_VALU) {
    if ((rc = fmbh_stream_create(&pl->aux))) return rc;
    if ((rc = fmbh_event_create_sync(&pl->ev_up))) return rc;
    if ((rc = fmbh_event_create_sync(&pl->ev_aux))) return rc;
    for (int l = 2; l <= L; l++)
      if ((rc = fmbh_event_create_sync(&pl->ev_m2l[l]))) return rc;
  }
  return fmbh_near_layout(pl, stream);
}

/* rotation blocks of every polar angle at order pl->Pmax, generated in FP64
 * and converted per opset, with the padded blocks and phase tables of the
 * matrix-core translations */
static int plan_rot(fmb_plan *pl, void *stream) {
  size_t b0 = pl->bytes;
  size_t hs = (size_t)fmb_rot_size(pl->Pmax) * pl->nbeta;
  double betas[FMB_NOFF], *d_H64 = NULL;
  for (int i = 0; i < pl->nbeta; i++) betas[i] = acos(pl->beta_cos[i]);
  int rc = fmbh_malloc((void **)&d_H64, sizeof(double) * hs);
  if (!rc) rc = fmbh_gen_H(betas, pl->nbeta, pl->Pmax, d_H64, stream);
  for (int i = 0; i < pl->nops && !rc; i++) {
    fmb_opset *o = &pl->ops[i];
    size_t bytes = o->csz / 2 * hs;
    rc = fmbh_malloc(&o->d_H, bytes);
    if (rc) break;
    pl->bytes += bytes;
    rc = fmbh_convert_real(o->prec, d_H64, hs, o->d_H, stream);
  }
  if (!rc) rc = fmbh_sync(stream);
  fmbh_free(d_H64);
  if (!rc) rc = fmbx_rcr_rot_setup(pl, stream);
  pl->bytes_rot += pl->bytes - b0;
  return rc;
}

/* operators of one precision from the FP64 coaxial originals */
static int build_opset(fmb_plan *pl, fmb_opset *o, fmb_c64 *const *d_sr64,
                       fmb_c64 *const *d_rr64, void *st) {
  const fmb_tree *t = pl->tree;
  int rc, L = t->lmax;
  for (int l = 2; l <= L; l++) {
    int p = pl->p[l];
    size_t cs = (size_t)fmb_coax_size(p);
    for (int d = 0; d < FMB_NDIST; d++) {
      /* M2L: L~ = w L, M = w M~  ->  diag(w) (S|R) diag(w) */
      if ((rc = fmbh_coax_scale(o->prec, d_sr64[l] + cs * d, p, pl->d_scale[l],
                                pl->d_scale[l], 0, 0,
                                (char *)o->d_Csr[l] + cs * d * o->csz, st)))
        return rc;
    }
    if (l > 2) {
      int P = pl->p[l] > pl->p[l - 1] ? pl->p[l] : pl->p[l - 1];
      /* M2M child l -> parent l-1: diag(1/w_{l-1}) (R|R) diag(w_l) */
      if ((rc = fmbh_coax_scale(o->prec, d_rr64[l], P, pl->d_scale[l - 1],
                                pl->d_scale[l], 1, 0, o->d_Cm2m[l], st)))
        return rc;
      /* L2L parent l-1 -> child l: diag(w_l) (R|R) diag(1/w_{l-1}) */
      if ((rc = fmbh_coax_scale(o->prec, d_rr64[l], P, pl->d_scale[l],
                                pl->d_scale[l - 1], 0, 1, o->d_Cl2l[l], st)))
        return rc;
    }
  }
  int op = (int)(o - pl->ops);
  if ((rc = fmbh_bases(pl, op, pl->d_robin, st))) return rc;
  if ((rc = fmbh_near_assemble(pl, op, pl->d_robin, st))) return rc;
  return fmbh_sync(st);
}

/* everything that depends on the wavenumber; the rotation tier is rebuilt
 * first when an order exceeds the current Pmax */
static int plan_freq(fmb_plan *pl, double k, const fmb_c64 *robin, void *stream) {
  const fmb_tree *t = pl->tree;
  const fmb_plan_params *prm = &pl->prm;
  int rc, L = t->lmax, P, p[FMB_MAXLEV] = {0};
  if ((rc = level_orders(pl, k, p, &P))) return rc;
  if (prm->k_max > k) {
    int pk[FMB_MAXLEV], Pk;
    if ((rc = level_orders(pl, prm->k_max, pk, &Pk))) return rc;
    if (Pk > P) P = Pk;
  }
  if (P > pl->Pmax) {
    plan_rot_free(pl);
    pl->Pmax = P;
    if ((rc = plan_rot(pl, stream))) return rc;
  }
  size_t b0 = pl->bytes;
  fmb_c64 *d_sr64[FMB_MAXLEV] = {0}, *d_rr64[FMB_MAXLEV] = {0};
  pl->k = k;
  if (prm->alpha_re == 0.0 && prm->alpha_im == 0.0) {
    pl->alpha_re = 0.0;
    pl->alpha_im = 1.0 / k; /* alpha = i/k */
  } else {
    pl->alpha_re = prm->alpha_re;
    pl->alpha_im = prm->alpha_im;
  }
  memcpy(pl->p, p, sizeof(p));
  /* scale factors w_n = min(1, 1/(x |h_n(x)|)), x = k * box radius */
  for (int l = 2; l <= L; l++) {
    double x = k * t->D * ldexp(1.0, -l) * sqrt(3.0) / 2.0;
    int n1 = pl->Pmax + 1;
    pl->h_scale[l] = (double *)malloc(sizeof(double) * n1);
    if (!pl->h_scale[l]) {
      rc = fmb_set_error(FMB_ERR_NOMEM, "scale factors");
      goto out;
    }
    fmbx_abs_hn(n1 - 1, x, pl->h_scale[l]);
    for (int n = 0; n < n1; n++) {
      double w = 1.0 / (x * pl->h_scale[l][n]);
      if (!(w > 1e-280)) w = 1e-280;
      pl->h_scale[l][n] = w < 1.0 ? w : 1.0;
    }
    if ((rc = fmbh_malloc((void **)&pl->d_scale[l], sizeof(double) * n1))) goto out;
    if ((rc = fmbh_h2d(pl->d_scale[l], pl->h_scale[l], sizeof(double) * n1, stream)))
      goto out;
    if ((rc = fmbh_mmajor_table(pl->p[l], &pl->d_tab_mmajor[l], stream))) goto out;
  }
  /* Robin coefficients in Morton order */
  if (robin) {
    int N = t->nelem;
    fmb_c64 *h = (fmb_c64 *)malloc(sizeof(fmb_c64) * N),
            *hs = (fmb_c64 *)malloc(sizeof(fmb_c64) * N);
    int *perm = (int *)malloc(sizeof(int) * N);
    rc = (h && hs && perm) ? FMB_OK : fmb_set_error(FMB_ERR_NOMEM, "Robin coefficients");
    if (!rc) rc = fmbh_d2h(h, robin, sizeof(fmb_c64) * N, stream);
    if (!rc) rc = fmbh_d2h(perm, t->d_perm, sizeof(int) * N, stream);
    if (!rc)
      for (int i = 0; i < N; i++) hs[i] = h[perm[i]];
    if (!rc) rc = fmbh_malloc((void **)&pl->d_robin, sizeof(fmb_c64) * N);
    if (!rc) rc = fmbh_h2d(pl->d_robin, hs, sizeof(fmb_c64) * N, stream);
    if (!rc) rc = fmbh_sync(stream);
    free(h);
    free(hs);
    free(perm);
    if (rc) goto out;
  }
  /* FP64 coaxial originals at this wavenumber */
  for (int l = 2; l <= L; l++) {
    double side = t->D * ldexp(1.0, -l), tv[FMB_NDIST];
    for (int d = 0; d < FMB_NDIST; d++) tv[d] = side * sqrt((double)k_dist2[d]);
    size_t cs = (size_t)fmb_coax_size(pl->p[l]) * FMB_NDIST;
    if ((rc = fmbh_malloc((void **)&d_sr64[l], sizeof(fmb_c64) * cs))) goto out;
    if ((rc = fmbh_gen_coax(FMB_COAX_SR, k, tv, FMB_NDIST, pl->p[l], d_sr64[l], stream)))
      goto out;
    if (l > 2) {
      int Pc = pl->p[l] > pl->p[l - 1] ? pl->p[l] : pl->p[l - 1];
      double r = side * sqrt(3.0) / 2.0;
      if ((rc = fmbh_malloc((void **)&d_rr64[l], sizeof(fmb_c64) * fmb_coax_size(Pc))))
        goto out;
      if ((rc = fmbh_gen_coax(FMB_COAX_RR, k, &r, 1, Pc, d_rr64[l], stream))) goto out;
    }
  }
  for (int i = 0; i < pl->nops; i++) {
    if ((rc = opset_alloc(pl, &pl->ops[i]))) goto out;
    if ((rc = build_opset(pl, &pl->ops[i], d_sr64, d_rr64, stream))) goto out;
  }
  rc = fmbx_rcr_setup(pl, stream);
out:
  for (int l = 0; l < FMB_MAXLEV; l++) {
    fmbh_free(d_sr64[l]);
    fmbh_free(d_rr64[l]);
  }
  pl->bytes_freq += pl->bytes - b0;
  if (!rc) pl->ready = 1;
  return rc;
}

int fmb_plan_create(const fmb_tree *t, const fmb_plan_params *prm,
                    const fmb_c64 *robin, void *stream, fmb_plan **out) {
  if (!t || !prm || !out || !(prm->k > 0) || prm->nrhs_max < 1 ||
      (prm->precision != FMB_FP64 && prm->precision != FMB_MIXED) ||
      !(prm->k_max >= 0))
    return fmb_set_error(FMB_ERR_ARG, "fmb_plan_create: invalid parameters");
  if (prm->nedge < 1 || prm->nedge > 256 || prm->near_ratio < 0)
    return fmb_set_error(FMB_ERR_ARG, "fmb_plan_create: quadrature parameters");
  if (prm->rcr_path != FMB_RCR_AUTO && prm->rcr_path != FMB_RCR_VALU)
    return fmb_set_error(FMB_ERR_ARG, "fmb_plan_create: rcr_path %d", prm->rcr_path);
  if (prm->waves != 0 && prm->waves != 4 && prm->waves != 8)
    return fmb_set_error(FMB_ERR_ARG, "fmb_plan_create: waves %d (0, 4 or 8)", prm->waves);
  if (prm->rcr_mode < FMB_RCR_MODE_RULE || prm->rcr_mode > FMB_RCR_MODE_LDSO)
    return fmb_set_error(FMB_ERR_ARG, "fmb_plan_create: rcr_mode %d", prm->rcr_mode);
  if (prm->m2l_tile < 0 || prm->m2l_tile_min_mb < 0)
    return fmb_set_error(FMB_ERR_ARG, "fmb_plan_create: m2l_tile %d, m2l_tile_min_mb %d (>= 0)",
                         prm->m2l_tile, prm->m2l_tile_min_mb);
  if (prm->split_bf16 && (prm->precision != FMB_MIXED || prm->rcr_path != FMB_RCR_AUTO))
    return fmb_set_error(FMB_ERR_ARG,
                         "fmb_plan_create: split_bf16 needs an FMB_MIXED plan on the matrix cores");
  int rc;
  fmb_plan *pl = (fmb_plan *)calloc(1, sizeof(fmb_plan));
  if (!pl) return fmb_set_error(FMB_ERR_NOMEM, "fmb_plan_create");
  pl->tree = t;
  pl->prm = *prm;
  pl->stream = stream;
  pl->nrhs_max = prm->nrhs_max;
  pl->nops = prm->precision == FMB_MIXED ? 2 : 1;
  pl->work_prec = prm->precision;
  for (int i = 0; i < pl->nops; i++) {
    pl->ops[i].prec = i == 0 ? FMB_FP64 : FMB_MIXED;
    pl->ops[i].csz = i == 0 ? 16 : 8;
  }
  if ((rc = plan_geometry(pl, stream)) || (rc = plan_freq(pl, prm->k, robin, stream))) {
    fmb_plan_destroy(pl);
    return rc;
  }
  *out = pl;
  return FMB_OK;
}

int fmb_plan_retune(fmb_plan *pl, double k, const fmb_c64 *robin, void *stream) {
  if (!pl || !(k > 0))
    return fmb_set_error(FMB_ERR_ARG, "fmb_plan_retune: invalid arguments");
  pl->stream = stream;
  plan_freq_free(pl);
  int rc = plan_freq(pl, k, robin, stream);
  if (rc) plan_freq_free(pl);
  return rc;
}

int fmb_plan_wavenumber(const fmb_plan *pl, double *k) {
  if (!pl || !k) return fmb_set_error(FMB_ERR_ARG, "null plan");
  *k = pl->ready ? pl->k : 0.0;
  return FMB_OK;
}

int fmb_plan_orders(const fmb_plan *pl, int *p16) {
  if (!pl || !p16) return fmb_set_error(FMB_ERR_ARG, "null plan");
  for (int l = 0; l < 16; l++) p16[l] = l < FMB_MAXLEV ? pl->p[l] : 0;
  return FMB_OK;
}

int fmb_plan_memory(const fmb_plan *pl, double *bytes) {
  if (!pl || !bytes) return fmb_set_error(FMB_ERR_ARG, "null plan");
  *bytes = (double)pl->bytes;
  return FMB_OK;
}

int fmb_plan_set_work_precision(fmb_plan *pl, int prec) {
  if (!pl || (prec != FMB_FP64 && prec != FMB_MIXED) ||
      (prec == FMB_MIXED && pl->nops < 2))
    return fmb_set_error(FMB_ERR_ARG, "working precision %d not built in this plan", prec);
  pl->work_prec = prec;
  return FMB_OK;
}

int fmb_set_timing(fmb_plan *pl, int on) {
  if (!pl) return fmb_set_error(FMB_ERR_ARG, "null plan");
  pl->timing = on;
  return FMB_OK;
}

int fmb_last_timing(const fmb_plan *pl, double *phase7, double *m2l16) {
  if (!pl) return fmb_set_error(FMB_ERR_ARG, "null plan");
  if (phase7) memcpy(phase7, pl->t_phase, sizeof(double) * 7);
  if (m2l16)
    for (int l = 0; l < 16; l++) m2l16[l] = l < FMB_MAXLEV ? pl->t_m2l[l] : 0.0;
  return FMB_OK;
}

/* phase events: 0 start, 1 after gather+P2M, 2 after M2M, 3 after M2L/L2L,
 * 4 after L2P+near */
int fmb_matvec(fmb_plan *pl, int prec, const fmb_c64 *x, fmb_c64 *y, int nrhs,
               void *stream) {
  return fmb_matvec_parts(pl, prec, 3, x, y, nrhs, stream);
}

int fmb_matvec_parts(fmb_plan *pl, int prec, int parts, const fmb_c64 *x,
                     fmb_c64 *y, int nrhs, void *stream) {
  if (!pl || parts < 1 || parts > 3 || !x || !y || nrhs < 1 || nrhs > pl->nrhs_max)
    return fmb_set_error(FMB_ERR_ARG, "fmb_matvec: nrhs %d (max %d)", nrhs,
                         pl ? pl->nrhs_max : 0);
  if (!pl->ready) return fmb_set_error(FMB_ERR_ARG, "fmb_matvec: plan holds no operator");
  int op = (prec == FMB_MIXED && pl->nops == 2) ? 1 : 0;
  const fmb_tree *t = pl->tree;
  fmb_opset *o = &pl->ops[op];
  int rc, L = t->lmax;
  void *ev[4 + 2 * FMB_MAXLEV] = {0};
  int tm = pl->timing;
  if (tm)
    for (int i = 0; i < 4 + 2 * FMB_MAXLEV; i++)
      if ((rc = fmbh_event_create(&ev[i]))) return rc;
  if (tm) fmbh_event_record(ev[0], stream);
  if ((rc = fmbh_gather(pl, op, x, nrhs, stream))) goto out;
  for (int l = 2; l <= L; l++) {
    size_t bytes = (size_t)t->nbox[l] * nrhs * pl->p[l] * pl->p[l] * o->csz;
    if ((rc = fmbh_memset(o->d_L[l], 0, bytes, stream))) goto out;
    /* parents receive the M2M translations of their children as sums */
    if (l < L && (rc = fmbh_memset(o->d_M[l], 0, bytes, stream))) goto out;
  }
  if ((rc = fmbh_p2m(pl, op, nrhs, stream))) goto out;
  if (tm) fmbh_event_record(ev[1], stream);
  for (int l = L; l > 2; l--)
    if ((rc = fmbx_rcr_run(pl, op, 1, l, nrhs, stream))) goto out;
  if (tm) fmbh_event_record(ev[2], stream);
  /* inside a graph capture the levels stay on the capturing stream (see
   * fmb_gmres_batch: a second fork there is not used) */
  if (!pl->aux || pl->capturing) {
    for (int l = 2; l <= L; l++) {
      if (tm) fmbh_event_record(ev[4 + 2 * l], stream);
      if ((rc = fmbx_rcr_run(pl, op, 0, l, nrhs, stream))) goto out;
      if (tm) fmbh_event_record(ev[5 + 2 * l], stream);
      if (l > 2 && (rc = fmbx_rcr_run(pl, op, 2, l, nrhs, stream))) goto out;
    }
  } else {
    /* M2L levels need only the upward pass; L2L into level l reads L_{l-1}
     * complete (its M2L and L2L), and adds into L_l with atomics beside
     * that level's M2L.  The caller's stream runs the M2L levels L-2 and
     * L-1; the second stream runs the coarser M2L levels and the L2L chain
     * up to level L-1, waiting for a fine level's M2L before the L2L that
     * reads it.  The finest M2L (the longest launch) then runs alone on the
     * caller's stream, followed by the last L2L. */
    const int ls = L - 2;
    void *aux = pl->aux;
    if ((rc = fmbh_event_record(pl->ev_up, stream))) goto out;
    if ((rc = fmbh_stream_wait(aux, pl->ev_up))) goto out;
    for (int l = ls; l < L; l++) {
      if (tm) fmbh_event_record(ev[4 + 2 * l], stream);
      if ((rc = fmbx_rcr_run(pl, op, 0, l, nrhs, stream))) goto out;
      if (tm) fmbh_event_record(ev[5 + 2 * l], stream);
      if ((rc = fmbh_event_record(pl->ev_m2l[l], stream))) goto out;
    }
    for (int l = 2; l < L; l++) {
      if (l < ls) {
        if (tm) fmbh_event_record(ev[4 + 2 * l], aux);
        if ((rc = fmbx_rcr_run(pl, op, 0, l, nrhs, aux))) goto out;
        if (tm) fmbh_event_record(ev[5 + 2 * l], aux);
      }
      if (l > 2) {
        if (l - 1 >= ls && (rc = fmbh_stream_wait(aux, pl->ev_m2l[l - 1]))) goto out;
        if ((rc = fmbx_rcr_run(pl, op, 2, l, nrhs, aux))) goto out;
      }
    }
    if ((rc = fmbh_event_record(pl->ev_aux, aux))) goto out;
    if ((rc = fmbh_stream_wait(stream, pl->ev_aux))) goto out;
    if (tm) fmbh_event_record(ev[4 + 2 * L], stream);
    if ((rc = fmbx_rcr_run(pl, op, 0, L, nrhs, stream))) goto out;
    if (tm) fmbh_event_record(ev[5 + 2 * L], stream);
    if ((rc = fmbx_rcr_run(pl, op, 2, L, nrhs, stream))) goto out;
  }
  if (tm) fmbh_event_record(ev[3], stream);
  if ((rc = fmbh_l2p_near(pl, op, nrhs, parts, stream))) goto out;
  if ((rc = fmbh_scatter(pl, op, y, nrhs, stream))) goto out;
  if (tm) {
    void *e4;
    fmbh_event_create(&e4);
    fmbh_event_record(e4, stream);
    float ms;
    fmbh_event_elapsed(ev[0], ev[1], &ms);
    pl->t_phase[1] = ms;
    fmbh_event_elapsed(ev[1], ev[2], &ms);
    pl->t_phase[2] = ms;
    fmbh_event_elapsed(ev[2], ev[3], &ms);
    pl->t_phase[3] = ms;
    fmbh_event_elapsed(ev[3], e4, &ms);
    pl->t_phase[5] = ms;
    fmbh_event_elapsed(ev[0], e4, &ms);
    pl->t_phase[6] = ms;
    for (int l = 2; l <= L; l++) {
      fmbh_event_elapsed(ev[4 + 2 * l], ev[5 + 2 * l], &ms);
      pl->t_m2l[l] = ms;
    }
    fmbh_event_destroy(e4);
  }
out:
  if (tm)
    for (int i = 0; i < 4 + 2 * FMB_MAXLEV; i++) fmbh_event_destroy(ev[i]);
  return rc;
}

/* y = A^T x: L2P^T, L2L^T upward, M2L^T and M2M^T downward, near^T + P2M^T
 * (the forward pass run backwards with every operator transposed; the
 * coefficients carry a (-1)^n factor, see fmbx_rcr_run_t).  near / basis
 * replace the opset's near blocks and P2M basis when not NULL. */
int fmbx_matvec_t(fmb_plan *pl, int op, int parts, const void *near,
                  const void *basis, const fmb_c64 *x, fmb_c64 *y, int nrhs,
                  void *stream) {
  const fmb_tree *t = pl->tree;
  fmb_opset *o = &pl->ops[op];
  int rc, L = t->lmax;
  if ((rc = fmbh_gather(pl, op, x, nrhs, stream))) return rc;
  if (parts & 2) {
    for (int l = 2; l <= L; l++) {
      size_t bytes = (size_t)t->nbox[l] * nrhs * pl->p[l] * pl->p[l] * o->csz;
      if (l < L && (rc = fmbh_memset(o->d_L[l], 0, bytes, stream))) return rc;
      if ((rc = fmbh_memset(o->d_M[l], 0, bytes, stream))) return rc;
    }
    if ((rc = fmbh_l2p_t(pl, op, nrhs, stream))) return rc;
    for (int l = L; l > 2; l--)
      if ((rc = fmbx_rcr_run_t(pl, op, 2, l, nrhs, stream))) return rc;
    for (int l = 2; l <= L; l++) {
      if (l > 2 && (rc = fmbx_rcr_run_t(pl, op, 1, l, nrhs, stream))) return rc;
      if ((rc = fmbx_rcr_run_t(pl, op, 0, l, nrhs, stream))) return rc;
    }
  }
  if ((rc = fmbh_near_t(pl, op, near, basis, nrhs, parts, stream))) return rc;
  return fmbh_scatter(pl, op, y, nrhs, stream);
}

int fmb_matvec_transpose(fmb_plan *pl, int prec, int parts, const fmb_c64 *x,
                         fmb_c64 *y, int nrhs, void *stream) {
  if (!pl || !pl->ready || parts < 1 || parts > 3 || !x || !y || nrhs < 1 ||
      nrhs > pl->nrhs_max)
    return fmb_set_error(FMB_ERR_ARG, "fmb_matvec_transpose: invalid arguments");
  if (pl->prm.rcr_path == FMB_RCR_VALU)
    return fmb_set_error(FMB_ERR_ARG,
                         "fmb_matvec_transpose: needs rcr_path FMB_RCR_AUTO");
  int op = (prec == FMB_MIXED && pl->nops == 2) ? 1 : 0;
  return fmbx_matvec_t(pl, op, parts, NULL, NULL, x, y, nrhs, stream);
}

int fmb_rhs_plane_wave(const fmb_plan *pl, const double *dirs, int nrhs,
                       fmb_c64 *b, void *stream) {
  if (!pl || !pl->ready || !dirs || !b || nrhs < 1)
    return fmb_set_error(FMB_ERR_ARG, "fmb_rhs_plane_wave: invalid arguments");
  return fmbh_rhs_plane(pl, dirs, nrhs, b, stream);
}

int fmb_field_eval(const fmb_plan *pl, const fmb_c64 *x, int nrhs,
                   const double *pts, int npts, const double *dirs, fmb_c64 *out,
                   void *stream) {
  if (!pl || !pl->ready || !x || nrhs < 1 || !pts || npts < 1 || !out)
    return fmb_set_error(FMB_ERR_ARG, "fmb_field_eval: invalid arguments");
  double *d_pts = NULL, *d_dirs = NULL;
  int rc = fmbh_malloc((void **)&d_pts, sizeof(double) * 3 * (size_t)npts);
  if (!rc) rc = fmbh_h2d(d_pts, pts, sizeof(double) * 3 * (size_t)npts, stream);
  if (!rc && dirs) rc = fmbh_malloc((void **)&d_dirs, sizeof(double) * 3 * (size_t)nrhs);
  if (!rc && dirs) rc = fmbh_h2d(d_dirs, dirs, sizeof(double) * 3 * (size_t)nrhs, stream);
  if (!rc) rc = fmbh_field(pl, x, nrhs, d_pts, npts, d_dirs, out, stream);
  if (!rc) rc = fmbh_sync(stream);
  fmbh_free(d_pts);
  fmbh_free(d_dirs);
  return rc;
}

typedef struct field_key {
  unsigned long long key;
  int idx;
} field_key;

static int cmp_field_key(const void *a, const void *b) {
  const field_key *x = (const field_key *)a, *y = (const field_key *)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->idx - y->idx;
}

int fmb_field_eval_fmm(fmb_plan *pl, const fmb_c64 *x, int nrhs,
                       const double *pts, int npts, const double *dirs,
                       fmb_c64 *out, void *stream) {
  if (!pl || !pl->ready || !x || nrhs < 1 || !pts || npts < 1 || !out)
    return fmb_set_error(FMB_ERR_ARG, "fmb_field_eval_fmm: invalid arguments");
  for (size_t i = 0; i < 3 * (size_t)npts; i++)
    if (!isfinite(pts[i]))
      return fmb_set_error(FMB_ERR_ARG, "fmb_field_eval_fmm: point %zu is not finite", i / 3);
  const fmb_tree *t = pl->tree;
  const int L = t->lmax, N = t->nelem;
  fmb_opset *o = &pl->ops[0];
  /* points in Morton order on a grid of 1/1024 of the cube over three cube
   * widths (points beyond clamp to its faces), so that the workgroups of
   * the box-major levels take compact groups of points */
  field_key *fk = (field_key *)malloc(sizeof(field_key) * (size_t)npts);
  double *spts = (double *)malloc(sizeof(double) * 3 * (size_t)npts);
  int *qperm = (int *)malloc(sizeof(int) * (size_t)npts);
  if (!fk || !spts || !qperm) {
    free(fk);
    free(spts);
    free(qperm);
    return fmb_set_error(FMB_ERR_NOMEM, "fmb_field_eval_fmm: point order");
  }
  for (int i = 0; i < npts; i++) {
    unsigned g[3];
    for (int c = 0; c < 3; c++) {
      double v = floor((pts[3 * (size_t)i + c] - t->pmin[c]) / t->D * 1024.0) + 1024.0;
      g[c] = v < 0.0 ? 0u : (v > 3071.0 ? 3071u : (unsigned)v);
    }
    unsigned long long key = 0;
    for (int b = 11; b >= 0; b--)
      key = key * 8 + 4 * ((g[0] >> b) & 1) + 2 * ((g[1] >> b) & 1) + ((g[2] >> b) & 1);
    fk[i].key = key;
    fk[i].idx = i;
  }
  qsort(fk, (size_t)npts, sizeof(field_key), cmp_field_key);
  for (int i = 0; i < npts; i++) {
    qperm[i] = fk[i].idx;
    for (int c = 0; c < 3; c++) spts[3 * (size_t)i + c] = pts[3 * (size_t)fk[i].idx + c];
  }
  free(fk);
  double *d_pts = NULL, *d_dirs = NULL;
  int *d_qperm = NULL;
  int rc = fmbh_malloc((void **)&d_pts, sizeof(double) * 3 * (size_t)npts);
  if (!rc) rc = fmbh_h2d(d_pts, spts, sizeof(double) * 3 * (size_t)npts, stream);
  if (!rc) rc = fmbh_malloc((void **)&d_qperm, sizeof(int) * (size_t)npts);
  if (!rc) rc = fmbh_h2d(d_qperm, qperm, sizeof(int) * (size_t)npts, stream);
  if (!rc && dirs) rc = fmbh_malloc((void **)&d_dirs, sizeof(double) * 3 * (size_t)nrhs);
  if (!rc && dirs) rc = fmbh_h2d(d_dirs, dirs, sizeof(double) * 3 * (size_t)nrhs, stream);
  for (int c0 = 0; c0 < nrhs && !rc; c0 += pl->nrhs_max) {
    const int nc = nrhs - c0 < pl->nrhs_max ? nrhs - c0 : pl->nrhs_max;
    const fmb_c64 *xc = x + (size_t)c0 * N;
    if ((rc = fmbh_gather(pl, 0, xc, nc, stream))) break;
    for (int l = 2; l < L && !rc; l++)
      rc = fmbh_memset(o->d_M[l], 0, (size_t)t->nbox[l] * nc * pl->p[l] * pl->p[l] * o->csz,
                       stream);
    if (!rc) rc = fmbh_p2m(pl, 0, nc, stream);
    for (int l = L; l > 2 && !rc; l--) rc = fmbx_rcr_run(pl, 0, 1, l, nc, stream);
    if (!rc)
      rc = fmbh_field_tree(pl, 0, xc, nc, d_pts, d_qperm, npts,
                           d_dirs ? d_dirs + 3 * c0 : NULL, out + (size_t)c0 * npts, stream);
  }
  /* the passes finish before their inputs are freed */
  int rs = fmbh_sync(stream);
  if (!rc) rc = rs;
  free(spts);
  free(qperm);
  fmbh_free(d_pts);
  fmbh_free(d_qperm);
  fmbh_free(d_dirs);
  return rc;
}

int fmb_sweep_guess(const fmb_plan *pl, double k_from, const fmb_c64 *x_from,
                    const double *dirs, int nrhs, fmb_c64 *x0, void *stream) {
  if (!pl || !pl->ready || !x_from || !x0 || nrhs < 1 || !(k_from > 0))
    return fmb_set_error(FMB_ERR_ARG, "fmb_sweep_guess: invalid arguments");
  return fmbh_sweep_guess(pl, pl->k - k_from, x_from, dirs, nrhs, x0, stream);
}
