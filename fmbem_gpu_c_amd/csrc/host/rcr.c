/* rcr.c -- host planning and dispatch of the rotate-coaxial-rotate
 * translations on the matrix cores (csrc/hip/rcr_kernels.h, rcr_mfma.hip).
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
 * workgroup's waves by longest-processing-time first.
 *
 * The FP32 / FP64 rotation blocks run in even/odd coordinates
 * (csrc/hip/rcr_mfma.hip): degree n's (2n+1)^2 block becomes an even block
 * of n + 1 rows and an odd block of n rows, about half the products, in one
 * unit per degree.  The split-BF16 rotations keep the standard blocks. */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/file.h>
#include <sys/stat.h>

#include "../fmb_internal.h"
#include "../hip/fmb_hip.h"
#include "host.h"

#define RCR_CHUNK 16
#define RCR_WAVES_MAX 8
#define RCR_GWS_GRID 512

typedef struct {
  int beta, dist;
  double phi;
} rcr_info_t; /* layout of the device OffInfo */

typedef struct {
  int code, cost, aoff; /* aoff: offset of the unit's padded operator block */
  int w;                /* wave the schedule gives it */
} unit_t;

static int by_cost(const void *a, const void *b) {
  const unit_t *x = (const unit_t *)a, *y = (const unit_t *)b;
  if (x->cost != y->cost) return y->cost - x->cost;
  return x->code - y->code;
}

/* units in non-increasing cost order, each to the least-loaded of nw
 * waves; writes the waves' lists back to back as {code, aoff} pairs into
 * out (entry q at out[2q]) and their start entries (nw + 1 values,
 * beginning at base) into off */
static void lpt(unit_t *u, int n, int nw, int base, int *off, int *out) {
  int load[RCR_WAVES_MAX] = {0}, cnt[RCR_WAVES_MAX] = {0};
  qsort(u, n, sizeof(unit_t), by_cost);
  for (int i = 0; i < n; i++) {
    int w = 0;
    for (int q = 1; q < nw; q++)
      if (load[q] < load[w]) w = q;
    load[w] += u[i].cost;
    u[i].w = w;
    cnt[w]++;
  }
  off[0] = base;
  for (int w = 0; w < nw; w++) off[w + 1] = off[w] + cnt[w];
  int fill[RCR_WAVES_MAX];
  for (int w = 0; w < nw; w++) fill[w] = off[w];
  for (int i = 0; i < n; i++) {
    int q = fill[u[i].w]++;
    out[2 * q] = u[i].code;
    out[2 * q + 1] = u[i].aoff;
  }
}

static int tiles(int s) { return (s + 15) / 16; }
static int ksteps(int s) { return (s + 3) / 4; }
static int group_tiles(int nt, int g, int mt, int split) {
  if (!split) return nt;
  return nt - g * mt < mt ? nt - g * mt : mt;
}

/* padded block sizes: rows to a multiple of 16, columns to a multiple of 4 */
static int pad_rows(int s) { return 16 * tiles(s); }
static int pad_cols(int s) { return 4 * ksteps(s); }

/* values of one padded even/odd rotation block of s rows (0 for s = 0) */
static int rot_block(int s) { return s ? pad_rows(s) * pad_cols(s) : 0; }

/* offsets of the padded rotation blocks of degrees n < np (np + 1 values):
 * degree n's even block (n + 1 rows), then its odd block (n rows) */
static void rot_offsets(int np, int *hoff) {
  hoff[0] = 0;
  for (int n = 0; n < np; n++) hoff[n + 1] = hoff[n] + rot_block(n + 1) + rot_block(n);
}

/* offsets of the split-BF16 (standard) rotation blocks n < np (np + 1
 * values): rows padded to 16, K to 32 */
static void rot_offsets_bf(int np, int *hb) {
  hb[0] = 0;
  for (int n = 0; n < np; n++) hb[n + 1] = hb[n] + pad_rows(2 * n + 1) * (((2 * n + 1) + 31) & ~31);
}

/* whether opset op may run mode's rotations in split-BF16 */
static int use_bf(const fmb_plan *pl, int op, int mode) {
  return pl->prm.split_bf16 && pl->ops[op].prec != FMB_FP64 &&
         (mode == FMBH_RCR_LDS || mode == FMBH_RCR_LDSP);  /* not LDSO */
}

/* offsets of the padded coaxial blocks |m| < min(pin, pout) (+1 values) */
static void coax_offsets(int pin, int pout, int *coff) {
  int mn = pin < pout ? pin : pout;
  coff[0] = 0;
  for (int am = 0; am < mn; am++)
    coff[am + 1] = coff[am] + pad_rows(pout - am) * pad_cols(pin - am);
}

/* rotation blocks of degrees n < np: code n*16 + g.  FP32 / FP64: degree
 * n's even (n + 1 rows) and odd (n rows) blocks together; split-BF16: the
 * standard block (2n + 1 rows). */
static int rot_units(int np, int split, int mt, const int *hoff, int bf, unit_t *u) {
  int k = 0;
  for (int n = 0; n < np; n++) {
    int s = bf ? 2 * n + 1 : n + 1, nt = tiles(s), ng = split ? (nt + mt - 1) / mt : 1;
    for (int g = 0; g < ng; g++) {
      u[k].code = n * 16 + g;
      u[k].cost = group_tiles(nt, g, mt, split) * ksteps(s);
      if (!bf) {
        /* the odd block's tiles in this row group */
        int to = split ? tiles(n) - g * mt : tiles(n);
        to = to < 0 ? 0 : (split && to > mt ? mt : to);
        u[k].cost += to * ksteps(n);
      }
      u[k].aoff = hoff[n];
      k++;
    }
  }
  return k;
}

/* coaxial blocks +-am, am < min(pin, pout), processed together (they share
 * the matrix C^{am}): code am*16 + g */
static int coax_units(int pin, int pout, int split, int mt, const int *coff,
                      unit_t *u) {
  int k = 0, mn = pin < pout ? pin : pout;
  for (int am = 0; am < mn; am++) {
    int so = pout - am, si = pin - am, sg = am ? 2 : 1;
    int nt = tiles(so), ng = split ? (nt + mt - 1) / mt : 1;
    for (int g = 0; g < ng; g++) {
      u[k].code = am * 16 + g;
      u[k].cost = 2 * sg * group_tiles(nt, g, mt, split) * ksteps(si);
      u[k].aoff = coff[am];
      k++;
    }
  }
  return k;
}

/* Schedule of one translation shape for nw waves: a header of three offset
 * tables (forward rotation, coaxial, back rotation), nw + 1 entry indices
 * each, giving every wave's units as entries [off[w], off[w+1]); entry q is
 * the int pair {code, operator offset} at [2q, 2q + 1] of the same array.
 * The header is 3 (nw + 1) ints rounded up to even. */
static int build_sched(int pin, int pout, int P, int mode, int mt, int nw, const int *hoff,
                       int bf, const int *coff, int **d_out, int *len, void *stream) {
  int split = mode == FMBH_RCR_GWS, rc;
  if (nw < 1 || nw > RCR_WAVES_MAX)
    return fmb_set_error(FMB_ERR_RANGE, "translation schedule for %d waves", nw);
  const int hdr = (3 * (nw + 1) + 1) & ~1;
  unit_t *u = (unit_t *)malloc(sizeof(unit_t) * 16 * (pin + P + pout));
  if (!u) return fmb_set_error(FMB_ERR_NOMEM, "translation schedule");
  int n1 = rot_units(pin, split, mt, hoff, bf, u);
  int nc = coax_units(pin, pout, split, mt, coff, u + n1);
  int n3 = rot_units(pout, split, mt, hoff, bf, u + n1 + nc);
  int tot = hdr + 2 * (n1 + nc + n3);
  int *h = (int *)calloc(tot, sizeof(int));
  if (!h) {
    free(u);
    return fmb_set_error(FMB_ERR_NOMEM, "translation schedule");
  }
  int base = hdr / 2;
  lpt(u, n1, nw, base, h, h);
  base += n1;
  lpt(u + n1, nc, nw, base, h + nw + 1, h);
  base += nc;
  lpt(u + n1 + nc, n3, nw, base, h + 2 * (nw + 1), h);
  *len = tot;
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
  if (!c) return fmb_set_error(FMB_ERR_NOMEM, "translation chunks");
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

/* M2L chunks in tile-major order: pairs of target boxes [tile * T, (tile+1)
 * * T) for every offset, then the next tile; within a (tile, offset) group
 * chunks of at most RCR_CHUNK pairs.  start: the offset buckets of the
 * tree's offset-major pair list, whose targets are in box order inside each
 * bucket. */
static int build_chunks_tiled(const fmb_tree *t, int l, int T, fmb_rcr_tab *tb, void *stream) {
  const int *start = t->h_off_start[l];
  int np = start[FMB_NOFF] - start[0], nbox = t->nbox[l], rc;
  int ntile = (nbox + T - 1) / T;
  int *tg = (int *)malloc(sizeof(int) * (np > 0 ? np : 1));
  /* per (tile, bucket): chunks counted, then placed */
  int *cnt = (int *)calloc((size_t)ntile * FMB_NOFF + 1, sizeof(int));
  int *c = (int *)malloc(sizeof(int) * 4 * (np / RCR_CHUNK + (size_t)ntile * FMB_NOFF + 1));
  if (!tg || !cnt || !c) {
    free(tg);
    free(cnt);
    free(c);
    return fmb_set_error(FMB_ERR_NOMEM, "translation chunks");
  }
  rc = fmbh_d2h(tg, t->d_pair_tgt[l], sizeof(int) * np, stream);
  int n = 0;
  if (!rc) {
    /* group boundaries inside each bucket: runs of one tile */
    for (int b = 0; b < FMB_NOFF; b++)
      for (int s0 = start[b]; s0 < start[b + 1];) {
        int tile = tg[s0] / T, e = s0;
        while (e < start[b + 1] && tg[e] / T == tile) e++;
        cnt[(size_t)tile * FMB_NOFF + b + 1] += (e - s0 + RCR_CHUNK - 1) / RCR_CHUNK;
        s0 = e;
      }
    for (size_t i = 1; i <= (size_t)ntile * FMB_NOFF; i++) cnt[i] += cnt[i - 1];
    for (int b = 0; b < FMB_NOFF; b++)
      for (int s0 = start[b]; s0 < start[b + 1];) {
        int tile = tg[s0] / T, e = s0;
        while (e < start[b + 1] && tg[e] / T == tile) e++;
        int q = cnt[(size_t)tile * FMB_NOFF + b];
        for (int f = s0; f < e; f += RCR_CHUNK, q++) {
          c[4 * q] = b;
          c[4 * q + 1] = f;
          c[4 * q + 2] = f + RCR_CHUNK < e ? f + RCR_CHUNK : e;
          c[4 * q + 3] = 0;
        }
        cnt[(size_t)tile * FMB_NOFF + b] = q;
        s0 = e;
      }
    n = cnt[(size_t)ntile * FMB_NOFF];
  }
  tb->nchunks = n;
  if (!rc) rc = fmbh_malloc(&tb->d_chunks, sizeof(int) * 4 * (n > 0 ? n : 1));
  if (!rc && n) rc = fmbh_h2d(tb->d_chunks, c, sizeof(int) * 4 * n, stream);
  free(tg);
  free(cnt);
  free(c);
  return rc;
}

/* (parent, child) pairs of level l bucketed by the child's octant */
static int octant_pairs(const fmb_tree *t, int l, fmb_rcr_tab *tb, void *stream) {
  int nb = t->nbox[l], rc;
  int *key = (int *)malloc(sizeof(int) * nb), *par = (int *)malloc(sizeof(int) * nb);
  int *sp = (int *)malloc(sizeof(int) * nb), *sc = (int *)malloc(sizeof(int) * nb);
  int start[9] = {0};
  rc = (key && par && sp && sc) ? FMB_OK : fmb_set_error(FMB_ERR_NOMEM, "octant pairs");
  if (!rc) rc = fmbh_d2h(key, t->d_keys[l], sizeof(int) * nb, stream);
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

/* tile shapes, schedules and padded coaxial blocks of one translation;
 * C: the scaled coaxial matrices of each opset (ncoax of them, order P) */
static int plan_tab(fmb_plan *pl, fmb_rcr_tab *tb, int pin, int pout,
                    void *const *C, int ncoax, const int *hoff, const int *hboff,
                    void *stream) {
  int rc = FMB_OK, mn = pin < pout ? pin : pout;
  tb->pin = pin;
  tb->pout = pout;
  tb->P = pin > pout ? pin : pout;
  int *coff = (int *)malloc(sizeof(int) * (mn + 1));
  int *d_coff = NULL;
  if (!coff) return fmb_set_error(FMB_ERR_NOMEM, "coaxial offsets");
  coax_offsets(pin, pout, coff);
  tb->cpad_stride = (size_t)coff[mn];
  tb->hlen = hoff[tb->P];
  for (int op = 0; op < pl->nops && !rc; op++) {
    fmb_opset *o = &pl->ops[op];
    tb->mode[op] = 0;
    if (pl->prm.rcr_path == FMB_RCR_VALU) continue;
    if ((rc = fmbh_rcr_shape(tb->P, o->prec, &tb->mode[op], &tb->nt[op], &tb->mt[op])))
      break;
    tb->nw[op] = pl->prm.waves ? pl->prm.waves : 4;
    /* a requested workspace mode: GWS always fits; LDSP wherever the rule's
     * LDS batch fits (the same LDS) */
    if (pl->prm.rcr_mode == FMB_RCR_MODE_GWS && tb->mode[op] != FMBH_RCR_GWS) {
      tb->mode[op] = FMBH_RCR_GWS;
      tb->nt[op] = 2;
      tb->mt[op] = 4;
    } else if (pl->prm.rcr_mode == FMB_RCR_MODE_LDSP && tb->mode[op] == FMBH_RCR_LDS) {
      tb->mode[op] = FMBH_RCR_LDSP;
    }
    /* operators in LDS where batch and operators fit (the schedule bound of
     * fmbh_rcr_shape) */
    if ((pl->prm.ldso || pl->prm.rcr_mode == FMB_RCR_MODE_LDSO) && tb->mode[op] != FMBH_RCR_GWS)
      for (int nt = 2; nt >= 1; nt--)
        if (fmbh_rcr_ldso_bytes(tb->P, o->prec, nt, 8 * tb->P + 16, tb->hlen,
                                tb->cpad_stride) <= 160 * 1024) {
          tb->mode[op] = FMBH_RCR_LDSO;
          tb->nt[op] = nt;
          break;
        }
    tb->bf[op] = use_bf(pl, op, tb->mode[op]);
    /* split-BF16 rotations keep the standard blocks: whole 2P - 1 rows */
    if (tb->bf[op] && fmbh_rcr_mt_lds(tb->P, 1)) tb->mt[op] = fmbh_rcr_mt_lds(tb->P, 1);
    if ((rc = build_sched(pin, pout, tb->P, tb->mode[op], tb->mt[op], tb->nw[op],
                          tb->bf[op] ? hboff : hoff, tb->bf[op], coff, &tb->d_sched[op],
                          &tb->sched_len[op], stream)))
      break;
    if (!d_coff) {
      if ((rc = fmbh_malloc((void **)&d_coff, sizeof(int) * (mn + 1)))) break;
      if ((rc = fmbh_h2d(d_coff, coff, sizeof(int) * (mn + 1), stream))) break;
    }
    size_t bytes = o->csz * tb->cpad_stride * ncoax;
    if ((rc = fmbh_malloc(&tb->d_Cpad[op], bytes))) break;
    pl->bytes += bytes;
    if ((rc = fmbh_pad_coax(o->prec, C[op], ncoax, tb->P, pin, pout, d_coff,
                            tb->cpad_stride, tb->d_Cpad[op], stream)))
      break;
  }
  if (!rc) rc = fmbh_sync(stream);
  fmbh_free(d_coff);
  free(coff);
  return rc;
}

int fmbx_rcr_rot_setup(fmb_plan *pl, void *stream) {
  if (pl->prm.rcr_path == FMB_RCR_VALU) return FMB_OK;
  int rc = FMB_OK;
  /* e^{i m phi} of every bucket, m < Pmax, per opset */
  int nph[3] = {FMB_NOFF, 8, 8};
  for (int q = 0; q < 3 && !rc; q++) {
    double *ph = (double *)malloc(sizeof(double) * 2 * nph[q] * pl->Pmax);
    for (int b = 0; b < nph[q]; b++) {
      double phi = q == 0 ? pl->off_phi[b] : (q == 1 ? pl->m2m_phi[b] : pl->l2l_phi[b]);
      for (int m = 0; m < pl->Pmax; m++) {
        ph[2 * ((size_t)b * pl->Pmax + m)] = cos(m * phi);
        ph[2 * ((size_t)b * pl->Pmax + m) + 1] = sin(m * phi);
      }
    }
    /* cj = 0: e^{i m phi}; cj = 1: e^{-i m phi} for the transposed pass */
    size_t n = (size_t)2 * nph[q] * pl->Pmax;
    for (int cj = 0; cj < 2 && !rc; cj++) {
      if (cj)
        for (size_t i = 1; i < n; i += 2) ph[i] = -ph[i];
      for (int op = 0; op < pl->nops && !rc; op++) {
        fmb_opset *o = &pl->ops[op];
        void **dst = cj ? &o->d_phase_c[q] : &o->d_phase[q];
        rc = fmbh_malloc(dst, o->csz / 2 * n);
        if (rc) break;
        if (o->prec == FMB_FP64) {
          rc = fmbh_h2d(*dst, ph, sizeof(double) * n, stream);
          if (!rc) rc = fmbh_sync(stream);
        } else {
          float *f = (float *)malloc(sizeof(float) * n);
          for (size_t i = 0; i < n; i++) f[i] = (float)ph[i];
          rc = fmbh_h2d(*dst, f, sizeof(float) * n, stream);
          if (!rc) rc = fmbh_sync(stream);
          free(f);
        }
      }
    }
    free(ph);
  }
  if (rc) return rc;
  /* padded rotation blocks of every polar angle, per opset */
  int *hoff = (int *)malloc(sizeof(int) * (pl->Pmax + 1)), *d_hoff = NULL;
  rot_offsets(pl->Pmax, hoff);
  rc = fmbh_malloc((void **)&d_hoff, sizeof(int) * (pl->Pmax + 1));
  if (!rc) rc = fmbh_h2d(d_hoff, hoff, sizeof(int) * (pl->Pmax + 1), stream);
  for (int op = 0; op < pl->nops && !rc; op++) {
    fmb_opset *o = &pl->ops[op];
    o->hpad_stride = (size_t)hoff[pl->Pmax];
    size_t bytes = o->csz / 2 * o->hpad_stride * pl->nbeta;
    rc = fmbh_malloc(&o->d_Hpad, bytes);
    if (!rc) {
      pl->bytes += bytes;
      /* from the FP64 blocks (opset 0), rounded once */
      rc = fmbh_pad_rot(o->prec, (const double *)pl->ops[0].d_H, pl->nbeta, pl->Pmax, d_hoff,
                        o->hpad_stride, o->d_Hpad, stream);
    }
    /* split-BF16 copies for the FP32 pass */
    if (!rc && use_bf(pl, op, FMBH_RCR_LDS)) {
      int *hb = (int *)malloc(sizeof(int) * (pl->Pmax + 1)), *d_hb = NULL;
      if (!hb) rc = fmb_set_error(FMB_ERR_NOMEM, "split-BF16 offsets");
      if (!rc) {
        rot_offsets_bf(pl->Pmax, hb);
        o->hb_half = (size_t)hb[pl->Pmax];
        size_t hbytes = 2 * sizeof(uint16_t) * o->hb_half * pl->nbeta;
        rc = fmbh_malloc((void **)&d_hb, sizeof(int) * (pl->Pmax + 1));
        if (!rc) rc = fmbh_h2d(d_hb, hb, sizeof(int) * (pl->Pmax + 1), stream);
        if (!rc) rc = fmbh_malloc(&o->d_Hb, hbytes);
        if (!rc) {
          pl->bytes += hbytes;
          rc = fmbh_split_rot((const double *)pl->ops[0].d_H, pl->nbeta, pl->Pmax, d_hb,
                              o->hb_half, o->d_Hb, stream);
        }
        if (!rc) rc = fmbh_sync(stream);
      }
      fmbh_free(d_hb);
      free(hb);
    }
  }
  if (!rc) rc = fmbh_sync(stream);
  fmbh_free(d_hoff);
  free(hoff);
  return rc;
}

/* global workspace of one opset: the largest GWS batch among its
 * translations (at least min_per_wg bytes per workgroup) */
static int size_gws(fmb_plan *pl, int op, size_t min_per_wg) {
  fmb_opset *o = &pl->ops[op];
  size_t need = min_per_wg;
  for (int kind = 0; kind < 3; kind++)
    for (int l = 2; l <= pl->tree->lmax; l++) {
      const fmb_rcr_tab *tb = &pl->rcr[kind][l];
      if (tb->mode[op] != FMBH_RCR_GWS) continue;
      size_t b = fmbh_rcr_gws_bytes(tb->P, o->prec, tb->nt[op]);
      if (b > need) need = b;
    }
  need *= RCR_GWS_GRID;
  o->gws_slot = need;
  /* a second slot for the launches on the plan's second stream */
  if (pl->aux) need *= 2;
  if (need == o->gws_bytes) return FMB_OK;
  fmbh_free(o->d_gws);
  o->d_gws = NULL;
  pl->bytes -= o->gws_bytes;
  o->gws_bytes = 0;
  o->gws_grid = 0;
  if (!need) return FMB_OK;
  int rc = fmbh_malloc(&o->d_gws, need);
  if (rc) return rc;
  o->gws_bytes = need;
  o->gws_grid = RCR_GWS_GRID;
  pl->bytes += need;
  return FMB_OK;
}

/* ---- autotuning of the tile shapes ----
 * Each translation (kind, level) of each opset is timed, on the plan's own
 * box pairs and buffers at nrhs_max right-hand sides, in every valid shape:
 * LDS mode with 8 or 16 vectors per batch (when the batch fits the 160 KB of
 * LDS), and global-workspace mode with 8 or 16 vectors and row groups of
 * 2, 3, 4 or 6 tiles, each with 4 or 8 waves per workgroup, and with
 * split_bf16 the LDS shapes once more with split-BF16 rotations.  The fastest
 * shape (a one-vector warm-up, then one timed run, HIP events on the plan's
 * stream) replaces the fixed rule of fmbh_rcr_shape. */
typedef struct {
  int mode, nt, mt, nw, bf;
} shape_t;

/* bfok: LDS shapes are also tried with split-BF16 rotations */
static int shape_candidates(int P, int bfok, shape_t *c) {
  int n = 0, mt = fmbh_rcr_mt_lds(P, 0);
  static const int mts[4] = {2, 3, 4, 6};
  for (int nw = 4; nw <= 8; nw += 4) {
    for (int bf = 0; bf <= bfok; bf++)
      for (int nt = 1; nt <= 2; nt++) {
        /* split-BF16 rotations keep the standard (2P - 1 row) blocks */
        const int m = fmbh_rcr_mt_lds(P, bf);
        if (m) {
          shape_t v = {FMBH_RCR_LDS, nt, m, nw, bf}, w = {FMBH_RCR_LDSP, nt, m, nw, bf};
          c[n++] = v;
          c[n++] = w;
        }
      }
    for (int nt = 1; nt <= 2; nt++)
      if (mt) {
        shape_t v = {FMBH_RCR_LDSO, nt, mt, nw, 0};
        c[n++] = v;
      }
    for (int nt = 1; nt <= 2; nt++)
      for (int i = 0; i < 4; i++) {
        shape_t v = {FMBH_RCR_GWS, nt, mts[i], nw, 0};
        c[n++] = v;
      }
  }
  return n;
}

/* one launch at one right-hand side (loads the kernel's code object and the
 * operator blocks into the caches), then one timed launch at nrhs_max */
static int time_run(fmb_plan *pl, int op, int kind, int l, void *stream, float *ms) {
  void *e0 = NULL, *e1 = NULL;
  int rc = fmbh_event_create(&e0);
  if (!rc) rc = fmbh_event_create(&e1);
  if (!rc) rc = fmbx_rcr_run(pl, op, kind, l, 1, stream);
  if (!rc) rc = fmbh_event_record(e0, stream);
  if (!rc) rc = fmbx_rcr_run(pl, op, kind, l, pl->nrhs_max, stream);
  if (!rc) rc = fmbh_event_record(e1, stream);
  if (!rc) rc = fmbh_sync(stream);
  if (!rc) rc = fmbh_event_elapsed(e0, e1, ms);
  fmbh_event_destroy(e0);
  fmbh_event_destroy(e1);
  return rc;
}

/* Shapes already tuned, keyed by what the timing depends on: device,
 * precision, kind, orders, chunk count and right-hand sides (a sweep that
 * revisits a level's order reuses its result).  The cache lives for the
 * process and in a text file shared by processes: FMBEM_TUNE_CACHE names it
 * (empty: no file), else $XDG_CACHE_HOME/fmbem/rcr_tiles.txt or
 * $HOME/.cache/fmbem/rcr_tiles.txt.  Each tuned shape is appended as one
 * line under an advisory lock; a process reads the file once, at its first
 * autotuned plan, and a later line for the same key wins. */
typedef struct {
  int prec, kind, pin, pout, nchunks, nrhs, bf;
  shape_t shape;
  float ms;
} tuned_t;
#define TUNE_CACHE 4096
#define TUNE_MAGIC "fmbem-rcr-tile-v2"
static tuned_t tune_cache[TUNE_CACHE];
static int tune_cached, tune_loaded;
static char tune_dev[64];
static pthread_mutex_t tune_lock = PTHREAD_MUTEX_INITIALIZER;

static int same_key(const tuned_t *c, const tuned_t *key) {
  return c->prec == key->prec && c->kind == key->kind && c->pin == key->pin &&
         c->pout == key->pout && c->nchunks == key->nchunks && c->nrhs == key->nrhs &&
         c->bf == key->bf;
}

/* caller holds tune_lock: replaces an entry of the same key or appends */
static void cache_store(const tuned_t *v) {
  for (int i = 0; i < tune_cached; i++)
    if (same_key(&tune_cache[i], v)) {
      tune_cache[i] = *v;
      return;
    }
  tune_cache[tune_cached < TUNE_CACHE ? tune_cached++ : TUNE_CACHE - 1] = *v;
}

/* the cache file's path, or 0 when there is none */
static int tune_path(char *buf, size_t n) {
  const char *e = getenv("FMBEM_TUNE_CACHE");
  if (e) {
    if (!*e || strlen(e) >= n) return 0;
    strcpy(buf, e);
    return 1;
  }
  const char *x = getenv("XDG_CACHE_HOME"), *h = getenv("HOME");
  int w;
  if (x && *x)
    w = snprintf(buf, n, "%s/fmbem/rcr_tiles.txt", x);
  else if (h && *h)
    w = snprintf(buf, n, "%s/.cache/fmbem/rcr_tiles.txt", h);
  else
    return 0;
  return w > 0 && (size_t)w < n;
}

/* creates the directories leading to path (mkdir -p of its dirname) */
static void make_parents(const char *path) {
  char d[1024];
  size_t n = strlen(path);
  if (n >= sizeof(d)) return;
  memcpy(d, path, n + 1);
  for (char *p = d + 1; *p; p++)
    if (*p == '/') {
      *p = 0;
      mkdir(d, 0755);
      *p = '/';
    }
}

/* reads the file's entries for this device (once per process) */
static void tune_load(void) {
  if (tune_loaded) return;
  tune_loaded = 1;
  if (fmbh_device_tag(tune_dev, sizeof(tune_dev))) tune_dev[0] = 0;
  char path[1024], line[512], magic[32], dev[64];
  if (!tune_path(path, sizeof(path))) return;
  FILE *f = fopen(path, "r");
  if (!f) return;
  flock(fileno(f), LOCK_SH);
  while (fgets(line, sizeof(line), f)) {
    tuned_t v;
    if (sscanf(line, "%31s %63s %d %d %d %d %d %d %d %d %d %d %d %d %f", magic, dev, &v.prec,
               &v.kind, &v.pin, &v.pout, &v.nchunks, &v.nrhs, &v.bf, &v.shape.mode,
               &v.shape.nt, &v.shape.mt, &v.shape.nw, &v.shape.bf, &v.ms) != 15)
      continue;
    if (strcmp(magic, TUNE_MAGIC) || strcmp(dev, tune_dev)) continue;
    /* only shapes this build can launch; split-BF16 only where allowed */
    if (v.shape.mode < FMBH_RCR_LDS || v.shape.mode > FMBH_RCR_LDSO || v.shape.nt < 1 ||
        v.shape.nt > 2 || (v.shape.nw != 4 && v.shape.nw != 8) || v.kind < 0 || v.kind > 2 ||
        (v.shape.mt != 2 && v.shape.mt != 3 && v.shape.mt != 4 && v.shape.mt != 6) ||
        (v.shape.bf && (!v.bf || v.shape.mode == FMBH_RCR_GWS ||
                        v.shape.mode == FMBH_RCR_LDSO)) || !(v.ms >= 0))
      continue;
    cache_store(&v);
  }
  flock(fileno(f), LOCK_UN);
  fclose(f);
}

static void tune_append(const tuned_t *v) {
  char path[1024];
  if (!tune_dev[0] || !tune_path(path, sizeof(path))) return;
  make_parents(path);
  FILE *f = fopen(path, "a");
  if (!f) return;
  flock(fileno(f), LOCK_EX);
  fprintf(f, "%s %s %d %d %d %d %d %d %d %d %d %d %d %d %.4f\n", TUNE_MAGIC, tune_dev,
          v->prec, v->kind, v->pin, v->pout, v->nchunks, v->nrhs, v->bf, v->shape.mode,
          v->shape.nt, v->shape.mt, v->shape.nw, v->shape.bf, v->ms);
  fflush(f);
  flock(fileno(f), LOCK_UN);
  fclose(f);
}

static int cache_find(const tuned_t *key, tuned_t *out) {
  int hit = 0;
  pthread_mutex_lock(&tune_lock);
  tune_load();
  for (int i = 0; i < tune_cached && !hit; i++)
    if (same_key(&tune_cache[i], key)) {
      *out = tune_cache[i];
      hit = 1;
    }
  pthread_mutex_unlock(&tune_lock);
  return hit;
}

static void cache_put(const tuned_t *v) {
  pthread_mutex_lock(&tune_lock);
  cache_store(v);
  tune_append(v);
  pthread_mutex_unlock(&tune_lock);
}

static int autotune(fmb_plan *pl, int op, void *stream) {
  const fmb_tree *t = pl->tree;
  fmb_opset *o = &pl->ops[op];
  int L = t->lmax, rc = FMB_OK;
  /* defined contents for the timing runs */
  for (int l = 2; l <= L && !rc; l++) {
    size_t bytes = (size_t)t->nbox[l] * pl->nrhs_max * pl->p[l] * pl->p[l] * o->csz;
    rc = fmbh_memset(o->d_M[l], 0, bytes, stream);
    if (!rc) rc = fmbh_memset(o->d_L[l], 0, bytes, stream);
  }
  int *hoff = (int *)malloc(sizeof(int) * (pl->Pmax + 1));
  int *hboff = (int *)malloc(sizeof(int) * (pl->Pmax + 1));
  if (!hoff || !hboff) {
    free(hoff);
    free(hboff);
    return fmb_set_error(FMB_ERR_NOMEM, "rotation offsets");
  }
  rot_offsets(pl->Pmax, hoff);
  rot_offsets_bf(pl->Pmax, hboff);
  for (int kind = 0; kind < 3 && !rc; kind++)
    for (int l = (kind ? 3 : 2); l <= L && !rc; l++) {
      fmb_rcr_tab *tb = &pl->rcr[kind][l];
      const fmb_rcr_tab *pairs = kind == 0 ? tb : &pl->rcr[1][l];
      if (!tb->mode[op] || !pairs->nchunks) continue;
      int mn = tb->pin < tb->pout ? tb->pin : tb->pout;
      int *coff = (int *)malloc(sizeof(int) * (mn + 1));
      if (!coff) {
        rc = fmb_set_error(FMB_ERR_NOMEM, "coaxial offsets");
        break;
      }
      coax_offsets(tb->pin, tb->pout, coff);
      tuned_t key = {o->prec, kind, tb->pin, tb->pout, pairs->nchunks, pl->nrhs_max,
                     use_bf(pl, op, FMBH_RCR_LDS), {0, 0, 0, 0, 0}, 0.0f}, hit;
      shape_t c[48];
      int nc;
      float best = 0.0f;
      int cached = cache_find(&key, &hit);
      if (cached) {
        /* only the cached shape, with its recorded time (no timing run) */
        c[0] = hit.shape;
        nc = 1;
        best = 1e30f;
      } else {
        rc = time_run(pl, op, kind, l, stream, &best);
        nc = shape_candidates(tb->P, use_bf(pl, op, FMBH_RCR_LDS), c);
      }
      for (int i = 0; i < nc && !rc; i++) {
        if (c[i].mode == tb->mode[op] && c[i].nt == tb->nt[op] && c[i].mt == tb->mt[op] &&
            c[i].nw == tb->nw[op] && c[i].bf == tb->bf[op])
          continue;
        int *sched = NULL, slen = 0;
        const int bf = c[i].bf;
        if ((rc = build_sched(tb->pin, tb->pout, tb->P, c[i].mode, c[i].mt, c[i].nw,
                              bf ? hboff : hoff, bf, coff, &sched, &slen, stream)))
          break;
        if (c[i].mode == FMBH_RCR_LDS || c[i].mode == FMBH_RCR_LDSP ||
            c[i].mode == FMBH_RCR_LDSO) {
          size_t b = c[i].mode == FMBH_RCR_LDSO
                         ? fmbh_rcr_ldso_bytes(tb->P, o->prec, c[i].nt, slen, tb->hlen,
                                               tb->cpad_stride)
                         : fmbh_rcr_lds_bytes(tb->P, o->prec, c[i].nt, slen);
          if (!b || b > 160 * 1024) {
            fmbh_free(sched);
            continue;
          }
        } else if ((rc = size_gws(pl, op, fmbh_rcr_gws_bytes(tb->P, o->prec, c[i].nt)))) {
          fmbh_free(sched);
          break;
        }
        shape_t prev = {tb->mode[op], tb->nt[op], tb->mt[op], tb->nw[op], tb->bf[op]};
        int *psched = tb->d_sched[op], plen = tb->sched_len[op];
        tb->mode[op] = c[i].mode;
        tb->nt[op] = c[i].nt;
        tb->mt[op] = c[i].mt;
        tb->nw[op] = c[i].nw;
        tb->bf[op] = bf;
        tb->d_sched[op] = sched;
        tb->sched_len[op] = slen;
        float ms = 0.0f;
        /* a cached shape is taken as it is; candidates are timed */
        if (!cached) rc = time_run(pl, op, kind, l, stream, &ms);
        if (!rc && ms < best) {
          best = ms;
          fmbh_free(psched);
        } else {
          tb->mode[op] = prev.mode;
          tb->nt[op] = prev.nt;
          tb->mt[op] = prev.mt;
          tb->nw[op] = prev.nw;
          tb->bf[op] = prev.bf;
          tb->d_sched[op] = psched;
          tb->sched_len[op] = plen;
          fmbh_free(sched);
        }
      }
      if (!rc && cached) best = hit.ms;
      tb->tuned_ms[op] = best;
      if (!rc && !cached) {
        key.shape.mode = tb->mode[op];
        key.shape.nt = tb->nt[op];
        key.shape.mt = tb->mt[op];
        key.shape.nw = tb->nw[op];
        key.shape.bf = tb->bf[op];
        key.ms = best;
        cache_put(&key);
      }
      free(coff);
    }
  free(hoff);
  free(hboff);
  if (!rc) rc = size_gws(pl, op, 0);
  return rc;
}

/* The tile-major M2L order keeps a tile's sources in its XCD's L2.  It pays
 * where the level's FP64 source coefficients (boxes x right-hand sides x
 * p^2 x 16 bytes) outgrow the GPU's 256 MB last-level (Infinity) cache, the
 * default m2l_tile_min_mb; below that the offset-major order's fuller chunks
 * win (measured: the level-6 and level-7 M2L of the 327,680-triangle sphere
 * at 9 right-hand sides 5-9% faster tile-major; the 20,480- and
 * 81,920-triangle spheres at one right-hand side 6-10% slower per matvec). */
static int m2l_tiled(const fmb_plan *pl, int l) {
  if (pl->prm.m2l_tile <= 0) return 0;
  const size_t p2 = (size_t)pl->p[l] * pl->p[l];
  const size_t bytes = (size_t)pl->tree->nbox[l] * pl->nrhs_max * p2 * 16;
  return bytes > ((size_t)(pl->prm.m2l_tile_min_mb > 0 ? pl->prm.m2l_tile_min_mb : 0) << 20);
}

int fmbx_rcr_setup(fmb_plan *pl, void *stream) {
  const fmb_tree *t = pl->tree;
  int L = t->lmax, rc = FMB_OK;
  for (int kind = 1; kind <= 2; kind++) {
    rcr_info_t info[8];
    for (int o = 0; o < 8; o++) {
      info[o].beta = kind == 1 ? pl->m2m_beta[o] : pl->l2l_beta[o];
      info[o].dist = 0;
      info[o].phi = kind == 1 ? pl->m2m_phi[o] : pl->l2l_phi[o];
    }
    if ((rc = fmbh_malloc(&pl->d_info_oct[kind - 1], sizeof(info)))) return rc;
    if ((rc = fmbh_h2d(pl->d_info_oct[kind - 1], info, sizeof(info), stream))) return rc;
    if ((rc = fmbh_sync(stream))) return rc;
  }
  if (pl->prm.rcr_path != FMB_RCR_VALU) {
    int *hoff = (int *)malloc(sizeof(int) * (pl->Pmax + 1));
    int *hboff = (int *)malloc(sizeof(int) * (pl->Pmax + 1));
    if (!hoff || !hboff) {
      free(hoff);
      free(hboff);
      return fmb_set_error(FMB_ERR_NOMEM, "rotation offsets");
    }
    rot_offsets(pl->Pmax, hoff);
    rot_offsets_bf(pl->Pmax, hboff);
    for (int l = 2; l <= L && !rc; l++) {
      fmb_rcr_tab *tb = &pl->rcr[0][l];
      void *C[2] = {pl->ops[0].d_Csr[l], pl->ops[1].d_Csr[l]};
      tb->tiled = m2l_tiled(pl, l);
      if (tb->tiled)
        rc = build_chunks_tiled(t, l, pl->prm.m2l_tile, tb, stream);
      else
        rc = build_chunks(t->h_off_start[l], FMB_NOFF, tb, stream);
      if (rc) break;
      rc = plan_tab(pl, tb, pl->p[l], pl->p[l], C, FMB_NDIST, hoff, hboff, stream);
    }
    for (int l = 3; l <= L && !rc; l++) {
      fmb_rcr_tab *up = &pl->rcr[1][l], *dn = &pl->rcr[2][l];
      void *Cu[2] = {pl->ops[0].d_Cm2m[l], pl->ops[1].d_Cm2m[l]};
      void *Cd[2] = {pl->ops[0].d_Cl2l[l], pl->ops[1].d_Cl2l[l]};
      if ((rc = octant_pairs(t, l, up, stream))) break;
      if ((rc = plan_tab(pl, up, pl->p[l], pl->p[l - 1], Cu, 1, hoff, hboff, stream))) break;
      rc = plan_tab(pl, dn, pl->p[l - 1], pl->p[l], Cd, 1, hoff, hboff, stream);
    }
    free(hoff);
    free(hboff);
    if (rc) return rc;
  }
  for (int op = 0; op < pl->nops; op++)
    if ((rc = size_gws(pl, op, 0))) return rc;
  if (pl->prm.autotune && pl->prm.rcr_path != FMB_RCR_VALU)
    for (int op = 0; op < pl->nops; op++)
      if ((rc = autotune(pl, op, stream))) return rc;
  return FMB_OK;
}

void fmbx_rcr_free(fmb_plan *pl) {
  for (int kind = 0; kind < 3; kind++)
    for (int l = 0; l < FMB_MAXLEV; l++) {
      fmb_rcr_tab *tb = &pl->rcr[kind][l];
      fmbh_free(tb->d_chunks);
      fmbh_free(tb->d_par);
      fmbh_free(tb->d_chi);
      for (int op = 0; op < 2; op++) {
        fmbh_free(tb->d_sched[op]);
        fmbh_free(tb->d_Cpad[op]);
      }
      memset(tb, 0, sizeof(*tb));
    }
  for (int i = 0; i < 2; i++) {
    fmbh_free(pl->d_info_oct[i]);
    pl->d_info_oct[i] = NULL;
  }
}

/* bytes of level l's coefficient array that a pass over nrhs vectors of
 * order p reaches ([box][nrhs][p^2]) */
static size_t coef_bytes(const fmb_plan *pl, const fmb_opset *o, int l, int p, int nrhs) {
  return (size_t)pl->tree->nbox[l] * nrhs * p * p * o->csz;
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
  j.nw = tb->nw[op];
  j.bf = tb->bf[op];
  j.Hb = o->d_Hb;
  j.hbstride = 2 * o->hb_half;
  j.hbhalf = o->hb_half;
  j.chunks = pairs->d_chunks;
  j.nchunks = pairs->nchunks;
  j.chunk_pairs = RCR_CHUNK;
  j.H = o->d_Hpad;
  j.hstride = o->hpad_stride;
  j.C = tb->d_Cpad[op];
  j.cstride = tb->cpad_stride;
  j.P = tb->P;
  j.pin = tb->pin;
  j.pout = tb->pout;
  j.sched = tb->d_sched[op];
  j.sched_len = tb->sched_len[op];
  j.hlen = tb->hlen;
  j.xcd = kind == 0 && tb->tiled && j.mode == FMBH_RCR_LDS;
  j.gws = (pl->aux && stream == pl->aux) ? (char *)o->d_gws + o->gws_slot : o->d_gws;
  j.gws_grid = o->gws_grid;
  j.phase = o->d_phase[kind];
  j.pstride = pl->Pmax;
  if (kind == 0) {
    j.info = pl->d_off_info;
    j.tgt = t->d_pair_tgt[l];
    j.src = t->d_pair_src[l];
    j.in = o->d_M[l];
    j.out = o->d_L[l];
  } else if (kind == 1) {
    j.info = pl->d_info_oct[0];
    j.tgt = pairs->d_par;
    j.src = pairs->d_chi;
    j.in = o->d_M[l];
    j.out = o->d_M[l - 1];
  } else {
    j.info = pl->d_info_oct[1];
    j.tgt = pairs->d_chi;
    j.src = pairs->d_par;
    j.in = o->d_L[l - 1];
    j.out = o->d_L[l];
  }
  /* source level: l (M2L, M2M) or l - 1 (L2L); target level: l or l - 1 */
  j.in_bytes = coef_bytes(pl, o, kind == 2 ? l - 1 : l, j.pin, nrhs);
  j.out_bytes = coef_bytes(pl, o, kind == 1 ? l - 1 : l, j.pout, nrhs);
  return fmbh_rcr(&j, o->prec, nrhs, stream);
}

/* The transposed translations.  The rotation blocks H_n(beta) are symmetric
 * and every coaxial |m| block satisfies C^T = P C P with P = diag((-1)^n),
 * so the transpose of T = Pout H C H Pin (Pin = e^{i m phi},
 * Pout = e^{-i m' phi}) is P (Pin H C' H Pout) P, where C' is C for M2L
 * and, for the scaled M2M / L2L blocks, the other direction's block:
 *   (diag(1/w_{l-1}) RR diag(w_l))^T = P diag(w_l) RR diag(1/w_{l-1}) P.
 * The transposed pass keeps its coefficients multiplied by P, so each
 * transposed translation is the forward kernel with conjugated phases:
 *   kind 0  M2L^T at l: L' (l) -> M' (l), pairs with source and target swapped;
 *   kind 1  M2M^T l-1 -> l: the L2L schedule and blocks, M2M's octant angles;
 *   kind 2  L2L^T l -> l-1: the M2M schedule and blocks, L2L's octant angles. */
int fmbx_rcr_run_t(const fmb_plan *pl, int op, int kind, int l, int nrhs,
                   void *stream) {
  const fmb_rcr_tab *tb = &pl->rcr[kind == 0 ? 0 : (kind == 1 ? 2 : 1)][l];
  if (!tb->mode[op])
    return fmb_set_error(FMB_ERR_ARG,
                         "transposed translations need the matrix-core kernels");
  const fmb_tree *t = pl->tree;
  const fmb_opset *o = &pl->ops[op];
  const fmb_rcr_tab *pairs = kind == 0 ? tb : &pl->rcr[1][l];
  fmbh_rcr_job j;
  memset(&j, 0, sizeof(j));
  j.mode = tb->mode[op];
  j.nt = tb->nt[op];
  j.mt = tb->mt[op];
  j.nw = tb->nw[op];
  j.bf = tb->bf[op];
  j.Hb = o->d_Hb;
  j.hbstride = 2 * o->hb_half;
  j.hbhalf = o->hb_half;
  j.chunks = pairs->d_chunks;
  j.nchunks = pairs->nchunks;
  j.chunk_pairs = RCR_CHUNK;
  j.H = o->d_Hpad;
  j.hstride = o->hpad_stride;
  j.C = tb->d_Cpad[op];
  j.cstride = tb->cpad_stride;
  j.P = tb->P;
  j.pin = tb->pin;
  j.pout = tb->pout;
  j.sched = tb->d_sched[op];
  j.sched_len = tb->sched_len[op];
  j.hlen = tb->hlen;
  j.xcd = kind == 0 && tb->tiled && j.mode == FMBH_RCR_LDS;
  j.gws = o->d_gws;
  j.gws_grid = o->gws_grid;
  j.phase = o->d_phase_c[kind];
  j.pstride = pl->Pmax;
  if (kind == 0) {
    j.info = pl->d_off_info;
    j.tgt = t->d_pair_src[l];
    j.src = t->d_pair_tgt[l];
    j.in = o->d_L[l];
    j.out = o->d_M[l];
  } else if (kind == 1) {
    j.info = pl->d_info_oct[0];
    j.tgt = pairs->d_chi;
    j.src = pairs->d_par;
    j.in = o->d_M[l - 1];
    j.out = o->d_M[l];
  } else {
    j.info = pl->d_info_oct[1];
    j.tgt = pairs->d_par;
    j.src = pairs->d_chi;
    j.in = o->d_L[l];
    j.out = o->d_L[l - 1];
  }
  /* source level: l (M2L^T, L2L^T) or l - 1 (M2M^T); target: l or l - 1 */
  j.in_bytes = coef_bytes(pl, o, kind == 1 ? l - 1 : l, j.pin, nrhs);
  j.out_bytes = coef_bytes(pl, o, kind == 2 ? l - 1 : l, j.pout, nrhs);
  return fmbh_rcr(&j, o->prec, nrhs, stream);
}

int fmb_plan_tile_shape(const fmb_plan *pl, int kind, int level, int prec, int *mode,
                        int *nt, int *mt, int *nw, int *bf, double *ms) {
  if (!pl || kind < 0 || kind > 2 || level < 2 || level > pl->tree->lmax ||
      (kind && level < 3))
    return fmb_set_error(FMB_ERR_ARG, "fmb_plan_tile_shape: invalid arguments");
  int op = (prec == FMB_MIXED && pl->nops == 2) ? 1 : 0;
  const fmb_rcr_tab *tb = &pl->rcr[kind][level];
  if (mode) *mode = tb->mode[op];
  if (nt) *nt = tb->nt[op];
  if (mt) *mt = tb->mt[op];
  if (nw) *nw = tb->nw[op];
  if (bf) *bf = tb->bf[op];
  if (ms) *ms = tb->tuned_ms[op];
  return FMB_OK;
}
