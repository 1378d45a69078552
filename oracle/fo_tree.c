/*
 * fo_tree.c -- Morton octree of the oracle: restatement of octree.c.
 *
 * TEST INFRASTRUCTURE ONLY (see fmb_oracle.h).
 *
 *   Morton numbering      octree.c:62-183 (dec2bin_frac, bitIntleave,
 *                         indArr2num, pnt2boxnum): x is the most significant
 *                         bit of each octal digit, coarse levels most
 *                         significant.
 *   boxCenter             octree.c:185-235: (t + 1/2) 2^-l in the unit cube
 *   neighbors / I4        octree.c:237-391, 517-543
 *   deterLmax             octree.c:635-679 (sorted level-10 keys, windows of
 *                         s+1 consecutive points)
 *   findBoundingCube      octree.c:681-747, with the z extents initialised
 *                         from z (the reference reads .y there,
 *                         octree.c:692-693; identical whenever the first
 *                         centroid's y lies inside the z range, as on every
 *                         shipped sphere)
 *   srcBoxes / FMMLevelSet / FMMLevelSetNumSR
 *                         octree.c:749-768, 890-919; sets are kept sorted
 *                         (the reference keeps first-appearance order;
 *                         compare as sets).
 *   truncNum              octree.c:777-796
 * Pinned by tests/test_oracle_tree.py: brute-force geometric definitions of
 * neighbours and interaction lists (|I4| = 189 for interior boxes,
 * 316 offsets, 15 distances), brute-force leaf depth (smallest level whose
 * fullest box holds <= s centroids), and the centroid bounding cube.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fmb_oracle.h"

int fo_trunc_num(double k, double eps, double sigma, double a) {
  double t0 = eps * pow(1.0 - 1.0 / sigma, 1.5);
  double t1 = log(t0) / log(sigma);
  double p_lo = 1.0 - t1;
  double ka = k * a;
  double c = pow(3.0 * log(1.0 / eps), 1.5) / 2.0;
  double p_hi = ka + c * pow(ka, 1.0 / 3.0);
  double p = ceil(pow(pow(p_lo, 4) + pow(p_hi, 4), 0.25));
  return (int)p;
}

static int grid_coord(double s, int l) {
  int n = 1 << l;
  int t = (int)floor(s * (double)n);
  if (t < 0) t = 0;
  if (t > n - 1) t = n - 1;
  return t;
}

static int interleave(int tx, int ty, int tz, int l) {
  int key = 0;
  for (int i = l - 1; i >= 0; i--)
    key = key * 8 + 4 * ((tx >> i) & 1) + 2 * ((ty >> i) & 1) + ((tz >> i) & 1);
  return key;
}

static void deinterleave(int key, int l, int *tx, int *ty, int *tz) {
  int x = 0, y = 0, z = 0;
  for (int i = 0; i < l; i++) {
    int d = (key >> (3 * i)) & 7;
    x |= ((d >> 2) & 1) << i;
    y |= ((d >> 1) & 1) << i;
    z |= (d & 1) << i;
  }
  *tx = x;
  *ty = y;
  *tz = z;
}

int fo_morton_key(const double *s, int l) {
  return interleave(grid_coord(s[0], l), grid_coord(s[1], l),
                    grid_coord(s[2], l), l);
}

void fo_box_center(int key, int l, double *c) {
  int tx, ty, tz;
  deinterleave(key, l, &tx, &ty, &tz);
  double h = ldexp(1.0, -l);
  c[0] = h * tx + 0.5 * h;
  c[1] = h * ty + 0.5 * h;
  c[2] = h * tz + 0.5 * h;
}

/* neighbours excluding the box itself, ascending key order */
static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

int fo_neighbors(int key, int l, int *nbr) {
  int tx, ty, tz, n = 0, N = 1 << l;
  deinterleave(key, l, &tx, &ty, &tz);
  for (int dx = -1; dx <= 1; dx++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dz = -1; dz <= 1; dz++) {
        int x = tx + dx, y = ty + dy, z = tz + dz;
        if (x < 0 || y < 0 || z < 0 || x >= N || y >= N || z >= N) continue;
        if (!dx && !dy && !dz) continue;
        nbr[n++] = interleave(x, y, z, l);
      }
  qsort(nbr, n, sizeof(int), cmp_int);
  return n;
}

/* I4 (octree.c:517-543): children of the parent's neighbourhood (parent
 * included) minus the box's own neighbourhood (box included). */
int fo_I4(int key, int l, int *set) {
  int pk = key >> 3;
  int pn[27], cn[27];
  int np = fo_neighbors(pk, l - 1, pn);
  pn[np++] = pk;
  int nc = fo_neighbors(key, l, cn);
  cn[nc++] = key;
  qsort(cn, nc, sizeof(int), cmp_int);
  int n = 0;
  for (int i = 0; i < np; i++)
    for (int c = 0; c < 8; c++) {
      int ch = pn[i] * 8 + c;
      if (bsearch(&ch, cn, nc, sizeof(int), cmp_int)) continue;
      set[n++] = ch;
    }
  qsort(set, n, sizeof(int), cmp_int);
  return n;
}

int fo_deter_lmax(const double *pts, int n, int s) {
  const int lav = 10;
  int *k = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int i = 0; i < n; i++) k[i] = fo_morton_key(pts + 3 * i, lav);
  qsort(k, n, sizeof(int), cmp_int);
  int lmax = 0;
  for (int i = 0, j = s; j < n; i++, j++) {
    int a = k[i], b = k[j], l = lav;
    if (a == b) {
      free(k);
      return 11; /* reference sentinel, octree.c:662-665 */
    }
    while (a != b) {
      l--;
      a >>= 3;
      b >>= 3;
    }
    if (l + 1 > lmax) lmax = l + 1;
  }
  free(k);
  return lmax;
}

int fo_bounding_cube(const double *pts, int n, double eps, double *pmin,
                     double *d) {
  double lo[3], hi[3];
  for (int c = 0; c < 3; c++) lo[c] = hi[c] = pts[c];
  for (int i = 1; i < n; i++)
    for (int c = 0; c < 3; c++) {
      double v = pts[3 * i + c];
      if (v < lo[c]) lo[c] = v;
      if (v > hi[c]) hi[c] = v;
    }
  double ext = hi[0] - lo[0];
  if (hi[1] - lo[1] < ext) ext = hi[1] - lo[1];
  if (hi[2] - lo[2] < ext) ext = hi[2] - lo[2];
  double e = 1e-6 * ext;
  if (eps < e) e = eps;
  double dm = 0.0, dd[3];
  for (int c = 0; c < 3; c++) {
    hi[c] += e;
    lo[c] -= e;
    dd[c] = hi[c] - lo[c];
    if (dd[c] > dm) dm = dd[c];
  }
  for (int c = 0; c < 3; c++) pmin[c] = lo[c] - 0.5 * (dm - dd[c]);
  *d = dm;
  return 0;
}

typedef struct {
  int key, idx;
} kv_t;
static int cmp_kv(const void *a, const void *b) {
  const kv_t *x = (const kv_t *)a, *y = (const kv_t *)b;
  if (x->key != y->key) return (x->key > y->key) - (x->key < y->key);
  return (x->idx > y->idx) - (x->idx < y->idx);
}

int fo_tree_build(const double *V, int nv, const int *F, int nf, int s,
                  int depth, fo_tree **out) {
  (void)nv;
  if (nf < 1) return -1;
  fo_tree *t = (fo_tree *)calloc(1, sizeof(fo_tree));
  double *ctr = (double *)malloc(sizeof(double) * 3 * nf);
  if (!t || !ctr) {
    free(t);
    free(ctr);
    return -1;
  }
  t->nelem = nf;
  t->lmin = 2;
  for (int i = 0; i < nf; i++) {
    const double *a = V + 3 * F[3 * i], *b = V + 3 * F[3 * i + 1],
                 *c = V + 3 * F[3 * i + 2];
    /* triCentroid_d (structs.cu:126-131) */
    for (int q = 0; q < 3; q++)
      ctr[3 * i + q] = a[q] + 2.0 / 3.0 * (0.5 * (b[q] + c[q]) - a[q]);
  }
  fo_bounding_cube(ctr, nf, 0.01, t->pmin, &t->D);
  for (int i = 0; i < nf; i++)
    for (int q = 0; q < 3; q++)
      ctr[3 * i + q] = (ctr[3 * i + q] - t->pmin[q]) / t->D;
  int L = depth > 0 ? depth : fo_deter_lmax(ctr, nf, s);
  if (L < 2) L = 2;
  if (L > 10) {
    free(ctr);
    free(t);
    return -1;
  }
  t->lmax = L;
  kv_t *kv = (kv_t *)malloc(sizeof(kv_t) * nf);
  /* elements ordered by their level-10 key, then by id (a refinement of the
   * leaf order used for the element permutation) */
  for (int i = 0; i < nf; i++) {
    kv[i].key = fo_morton_key(ctr + 3 * i, 10);
    kv[i].idx = i;
  }
  qsort(kv, nf, sizeof(kv_t), cmp_kv);
  for (int i = 0; i < nf; i++) kv[i].key >>= 3 * (10 - L);
  t->perm = (int *)malloc(sizeof(int) * nf);
  t->elem_box = (int *)malloc(sizeof(int) * nf);
  int nb = 0;
  int *lk = (int *)malloc(sizeof(int) * nf);
  t->box_start = (int *)malloc(sizeof(int) * nf);
  t->box_end = (int *)malloc(sizeof(int) * nf);
  for (int i = 0; i < nf; i++) {
    t->perm[i] = kv[i].idx;
    if (i == 0 || kv[i].key != kv[i - 1].key) {
      lk[nb] = kv[i].key;
      t->box_start[nb] = i;
      nb++;
    }
    t->box_end[nb - 1] = i + 1;
    t->elem_box[kv[i].idx] = nb - 1;
  }
  t->keys[L] = lk;
  t->nbox[L] = nb;
  for (int l = L - 1; l >= 2; l--) {
    int *ck = t->keys[l + 1], cn = t->nbox[l + 1];
    int *pk = (int *)malloc(sizeof(int) * cn);
    int np = 0;
    for (int i = 0; i < cn; i++) {
      int p = ck[i] >> 3;
      if (np == 0 || pk[np - 1] != p) pk[np++] = p;
    }
    t->keys[l] = pk;
    t->nbox[l] = np;
  }
  free(kv);
  free(ctr);
  *out = t;
  return 0;
}

void fo_tree_free(fo_tree *t) {
  if (!t) return;
  for (int l = 0; l < 16; l++) free(t->keys[l]);
  free(t->perm);
  free(t->elem_box);
  free(t->box_start);
  free(t->box_end);
  free(t);
}

int fo_tree_lmax(const fo_tree *t) { return t->lmax; }
int fo_tree_nbox(const fo_tree *t, int l) {
  return (l >= 2 && l <= t->lmax) ? t->nbox[l] : 0;
}
void fo_tree_keys(const fo_tree *t, int l, int *keys) {
  memcpy(keys, t->keys[l], sizeof(int) * t->nbox[l]);
}
void fo_tree_geom(const fo_tree *t, double *D, double *pmin) {
  *D = t->D;
  memcpy(pmin, t->pmin, sizeof(double) * 3);
}
void fo_tree_perm(const fo_tree *t, int *perm) {
  memcpy(perm, t->perm, sizeof(int) * t->nelem);
}

/* FMMLevelSetNumSR (octree.c:908-919): |I4(b) & level set| per box */
int fo_tree_num_sr(const fo_tree *t, int l, int *num) {
  int set[216];
  int total = 0;
  for (int i = 0; i < t->nbox[l]; i++) {
    int n = fo_I4(t->keys[l][i], l, set), c = 0;
    for (int j = 0; j < n; j++)
      if (bsearch(&set[j], t->keys[l], t->nbox[l], sizeof(int), cmp_int)) c++;
    num[i] = c;
    total += c;
  }
  return total;
}
