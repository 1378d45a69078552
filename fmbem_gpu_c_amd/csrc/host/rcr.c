/* rcr.c -- host planning and dispatch of the rotate-coaxial-rotate
 * translations on the matrix cores (csrc/hip/rcr_mfma.hip).
 *
 * Every translation kind groups its box pairs into buckets that share one
 * operator:
 *   M2L at level l: the 316 interaction-list offsets (srTransOriginArr /
 *     srTransDestArr, translation.cu:2847-2927), pairs already bucketed by
 *     the tree build;
 *   M2M l -> l-1 and L2L l-1 -> l: the child's octant (the 8 directions of
 *     genSSIdxArrs / genRRIdxArrs, translation.cu:2797-2846, 2928-2969),
 *     one (parent, child) list per level serving both.
 * Buckets are cut into chunks of at most RCR_CHUNK pairs.  The blocks of one
 * translation are split into units (whole blocks when the batch sits in LDS
 * and is updated in place; row groups otherwise) and assigned to the
 * workgroup's waves by longest-processing-time first. */
#include <stdlib.h>
#include <string.h>

#include "../fmb_internal.h"
#include "../hip/fmb_hip.h"
#include "host.h"

#define RCR_CHUNK 16
#define RCR_WAVES 4
#define RCR_GWS_GRID 512

typedef struct {
  int beta, dist;
  double phi;
} rcr_info_t; /* layout of the device OffInfo */

typedef struct {
  int code, cost;
} unit_t;

static int by_cost(const void *a, const void *b) {
  const unit_t *x = (const unit_t *)a, *y = (const unit_t *)b;
  if (x->cost != y->cost) return y->cost - x->cost;
  return x->code - y->code;
}

/* units in non-increasing cost order, each to the least-loaded wave;
 * out[w * n + i], -1 padded */
static void lpt(unit_t *u, int n, int *out) {
  int load[RCR_WAVES] = {0}, cnt[RCR_WAVES] = {0};
  qsort(u, n, sizeof(unit_t), by_cost);
  for (int i = 0; i < RCR_WAVES * n; i++) out[i] = -1;
  for (int i = 0; i < n; i++) {
    int w = 0;
    for (int q = 1; q < RCR_WAVES; q++)
      if (load[q] < load[w]) w = q;
    load[w] += u[i].cost;
    out[w * n + cnt[w]++] = u[i].code;
  }
}

static int tiles(int s) { return (s + 15) / 16; }
static int ksteps(int s) { return (s + 3) / 4; }
static int group_tiles(int nt, int g, int mt, int split) {
  if (!split) return nt;
  return nt - g * mt < mt ? nt - g * mt : mt;
}

/* rotation blocks n < np: code n*16 + g */
static int rot_units(int np, int split, int mt, unit_t *u) {
  int k = 0;
  for (int n = 0; n < np; n++) {
    int s = 2 * n + 1, nt = tiles(s), ng = split ? (nt + mt - 1) / mt : 1;
    for (int g = 0; g < ng; g++) {
      u[k].code = n * 16 + g;
      u[k].cost = group_tiles(nt, g, mt, split) * ksteps(s);
      k++;
    }
  }
  return k;
}

/* coaxial blocks |m| < min(pin, pout): code (m + P)*16 + g */
static int coax_units(int pin, int pout, int P, int split, int mt, unit_t *u) {
  int k = 0, mn = pin < pout ? pin : pout;
  for (int m = -(mn - 1); m <= mn - 1; m++) {
    int am = m < 0 ? -m : m, so = pout - am, si = pin - am;
    int nt = tiles(so), ng = split ? (nt + mt - 1) / mt : 1;
    for (int g = 0; g < ng; g++) {
      u[k].code = (m + P) * 16 + g;
      u[k].cost = 2 * group_tiles(nt, g, mt, split) * ksteps(si);
      k++;
    }
  }
  return k;
}

/* {w1, wc, w3} then [4][w1] forward rotation, [4][wc] coaxial,
 * [4][w3] back rotation */
static int build_sched(int pin, int pout, int P, int mode, int mt, int **d_out,
                       void *stream) {
  int split = mode == FMBH_RCR_GWS, rc;
  unit_t *u = (unit_t *)malloc(sizeof(unit_t) * 16 * (pin + 2 * P + pout));
  int n1 = rot_units(pin, split, mt, u);
  int nc = coax_units(pin, pout, P, split, mt, u + n1);
  int n3 = rot_units(pout, split, mt, u + n1 + nc);
  int tot = 3 + RCR_WAVES * (n1 + nc + n3);
  int *h = (int *)malloc(sizeof(int) * tot);
  h[0] = n1;
  h[1] = nc;
  h[2] = n3;
  lpt(u, n1, h + 3);
  lpt(u + n1, nc, h + 3 + RCR_WAVES * n1);
  lpt(u + n1 + nc, n3, h + 3 + RCR_WAVES * (n1 + nc));
  rc = fmbh_malloc((void **)d_out, sizeof(int) * tot);
  if (!rc) rc = fmbh_h2d(*d_out, h, sizeof(int) * tot, stream);
  free(u);
  free(h);
  return rc;
}

/* chunks of <= RCR_CHUNK pairs inside each bucket [start[b], start[b+1]) */
static int build_chunks(const int *start, int nbucket, fmb_rcr_tab *tb,
                        void *stream) {
  int npairs = start[nbucket] - start[0], n = 0, rc;
  int *c = (int *)malloc(sizeof(int) * 4 * (npairs / RCR_CHUNK + nbucket + 1));
  for (int b = 0; b < nbucket; b++)
    for (int s = start[b]; s < start[b + 1]; s += RCR_CHUNK) {
      int e = s + RCR_CHUNK < start[b + 1] ? s + RCR_CHUNK : start[b + 1];
      c[4 * n] = b;
      c[4 * n + 1] = s;
      c[4 * n + 2] = e;
      c[4 * n + 3] = 0;
      n++;
    }
  tb->nchunks = n;
  rc = fmbh_malloc(&tb->d_chunks, sizeof(int) * 4 * (n > 0 ? n : 1));
  if (!rc && n) rc = fmbh_h2d(tb->d_chunks, c, sizeof(int) * 4 * n, stream);
  free(c);
  return rc;
}

/* (parent, child) pairs of level l bucketed by the child's octant */
static int octant_pairs(const fmb_tree *t, int l, fmb_rcr_tab *tb, void *stream) {
  int nb = t->nbox[l], rc;
  int *key = (int *)malloc(sizeof(int) * nb), *par = (int *)malloc(sizeof(int) * nb);
  int *sp = (int *)malloc(sizeof(int) * nb), *sc = (int *)malloc(sizeof(int) * nb);
  int start[9] = {0};
  rc = fmbh_d2h(key, t->d_keys[l], sizeof(int) * nb, stream);
  if (!rc) rc = fmbh_d2h(par, t->d_parent[l], sizeof(int) * nb, stream);
  if (!rc) {
    for (int i = 0; i < nb; i++) start[(key[i] & 7) + 1]++;
    for (int o = 0; o < 8; o++) start[o + 1] += start[o];
    int fill[8];
    memcpy(fill, start, sizeof(fill));
    for (int i = 0; i < nb; i++) {
      int o = key[i] & 7, q = fill[o]++;
      sp[q] = par[i];
      sc[q] = i;
    }
    rc = fmbh_malloc((void **)&tb->d_par, sizeof(int) * nb);
  }
  if (!rc) rc = fmbh_malloc((void **)&tb->d_chi, sizeof(int) * nb);
  if (!rc) rc = fmbh_h2d(tb->d_par, sp, sizeof(int) * nb, stream);
  if (!rc) rc = fmbh_h2d(tb->d_chi, sc, sizeof(int) * nb, stream);
  if (!rc) rc = build_chunks(start, 8, tb, stream);
  free(key);
  free(par);
  free(sp);
  free(sc);
  return rc;
}

static int plan_tab(fmb_plan *pl, fmb_rcr_tab *tb, int pin, int pout, void *stream) {
  int rc;
  tb->pin = pin;
  tb->pout = pout;
  tb->P = pin > pout ? pin : pout;
  for (int op = 0; op < pl->nops; op++) {
    tb->mode[op] = 0;
    if (pl->prm.rcr_path == FMB_RCR_VALU) continue;
    if ((rc = fmbh_rcr_shape(tb->P, pl->ops[op].prec, &tb->mode[op], &tb->nt[op],
                             &tb->mt[op])))
      return rc;
    if ((rc = build_sched(pin, pout, tb->P, tb->mode[op], tb->mt[op],
                          &tb->d_sched[op], stream)))
      return rc;
  }
  return FMB_OK;
}

int fmbx_rcr_setup(fmb_plan *pl, void *stream) {
  const fmb_tree *t = pl->tree;
  int L = t->lmax, rc;
  for (int kind = 1; kind <= 2; kind++) {
    rcr_info_t info[8];
    for (int o = 0; o < 8; o++) {
      info[o].beta = kind == 1 ? pl->m2m_beta[o] : pl->l2l_beta[o];
      info[o].dist = 0;
      info[o].phi = kind == 1 ? pl->m2m_phi[o] : pl->l2l_phi[o];
    }
    if ((rc = fmbh_malloc(&pl->d_info_oct[kind - 1], sizeof(info)))) return rc;
    if ((rc = fmbh_h2d(pl->d_info_oct[kind - 1], info, sizeof(info), stream))) return rc;
  }
  for (int l = 2; l <= L; l++) {
    fmb_rcr_tab *tb = &pl->rcr[0][l];
    if ((rc = build_chunks(t->h_off_start[l], FMB_NOFF, tb, stream))) return rc;
    if ((rc = plan_tab(pl, tb, pl->p[l], pl->p[l], stream))) return rc;
  }
  for (int l = 3; l <= L; l++) {
    fmb_rcr_tab *up = &pl->rcr[1][l], *dn = &pl->rcr[2][l];
    if ((rc = octant_pairs(t, l, up, stream))) return rc;
    if ((rc = plan_tab(pl, up, pl->p[l], pl->p[l - 1], stream))) return rc;
    if ((rc = plan_tab(pl, dn, pl->p[l - 1], pl->p[l], stream))) return rc;
  }
  /* global workspace: the largest GWS batch of each opset */
  for (int op = 0; op < pl->nops; op++) {
    size_t need = 0;
    for (int kind = 0; kind < 3; kind++)
      for (int l = 2; l <= L; l++) {
        const fmb_rcr_tab *tb = &pl->rcr[kind][l];
        if (tb->mode[op] != FMBH_RCR_GWS) continue;
        size_t b = fmbh_rcr_gws_bytes(tb->P, pl->ops[op].prec, tb->nt[op]);
        if (b > need) need = b;
      }
    if (!need) continue;
    pl->ops[op].gws_grid = RCR_GWS_GRID;
    if ((rc = fmbh_malloc(&pl->ops[op].d_gws, need * RCR_GWS_GRID))) return rc;
    pl->bytes += need * RCR_GWS_GRID;
  }
  return FMB_OK;
}

void fmbx_rcr_free(fmb_plan *pl) {
  for (int kind = 0; kind < 3; kind++)
    for (int l = 0; l < FMB_MAXLEV; l++) {
      fmb_rcr_tab *tb = &pl->rcr[kind][l];
      fmbh_free(tb->d_chunks);
      fmbh_free(tb->d_par);
      fmbh_free(tb->d_chi);
      for (int op = 0; op < 2; op++) fmbh_free(tb->d_sched[op]);
      memset(tb, 0, sizeof(*tb));
    }
  for (int i = 0; i < 2; i++) {
    fmbh_free(pl->d_info_oct[i]);
    pl->d_info_oct[i] = NULL;
  }
}

/* kind 0: M2L at level l; 1: M2M level l -> l-1; 2: L2L level l-1 -> l */
int fmbx_rcr_run(const fmb_plan *pl, int op, int kind, int l, int nrhs,
                 void *stream) {
  const fmb_rcr_tab *tb = &pl->rcr[kind][l];
  if (!tb->mode[op]) {
    if (kind == 0) return fmbh_m2l(pl, op, l, nrhs, stream);
    if (kind == 1) return fmbh_m2m(pl, op, l, nrhs, stream);
    return fmbh_l2l(pl, op, l, nrhs, stream);
  }
  const fmb_tree *t = pl->tree;
  const fmb_opset *o = &pl->ops[op];
  const fmb_rcr_tab *pairs = kind == 0 ? tb : &pl->rcr[1][l];
  fmbh_rcr_job j;
  memset(&j, 0, sizeof(j));
  j.mode = tb->mode[op];
  j.nt = tb->nt[op];
  j.mt = tb->mt[op];
  j.chunks = pairs->d_chunks;
  j.nchunks = pairs->nchunks;
  j.chunk_pairs = RCR_CHUNK;
  j.H = o->d_H;
  j.hstride = (size_t)fmb_rot_size(pl->Pmax);
  j.P = tb->P;
  j.pin = tb->pin;
  j.pout = tb->pout;
  j.sched = tb->d_sched[op];
  j.gws = o->d_gws;
  j.gws_grid = o->gws_grid;
  if (kind == 0) {
    j.info = pl->d_off_info;
    j.tgt = t->d_pair_tgt[l];
    j.src = t->d_pair_src[l];
    j.C = o->d_Csr[l];
    j.cstride = (size_t)fmb_coax_size(tb->P);
    j.in = o->d_M[l];
    j.out = o->d_L[l];
  } else if (kind == 1) {
    j.info = pl->d_info_oct[0];
    j.tgt = pairs->d_par;
    j.src = pairs->d_chi;
    j.C = o->d_Cm2m[l];
    j.in = o->d_M[l];
    j.out = o->d_M[l - 1];
  } else {
    j.info = pl->d_info_oct[1];
    j.tgt = pairs->d_chi;
    j.src = pairs->d_par;
    j.C = o->d_Cl2l[l];
    j.in = o->d_L[l - 1];
    j.out = o->d_L[l];
  }
  return fmbh_rcr(&j, o->prec, nrhs, stream);
}
