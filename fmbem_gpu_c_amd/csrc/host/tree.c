/* tree.c -- octree set-up on the device: srcBoxes (octree.c:749-768),
 * deterLmax (octree.c:635-679), FMMLevelSet (octree.c:890-906), the
 * interaction lists of genOctree (translation.cu:2797-2969) and the leaf
 * neighbourhoods used by the near field. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../fmb_internal.h"
#include "../hip/fmb_hip.h"

/* findBoundingCube (octree.c:681-747) with the z extents taken from z */
static void bounding_cube(const double *lo0, const double *hi0, double *pmin,
                          double *D) {
  double lo[3], hi[3], d[3], dm = 0.0;
  for (int q = 0; q < 3; q++) {
    lo[q] = lo0[q];
    hi[q] = hi0[q];
  }
  double ext = hi[0] - lo[0];
  if (hi[1] - lo[1] < ext) ext = hi[1] - lo[1];
  if (hi[2] - lo[2] < ext) ext = hi[2] - lo[2];
  double e = 1e-6 * ext;
  if (0.01 < e) e = 0.01;
  for (int q = 0; q < 3; q++) {
    hi[q] += e;
    lo[q] -= e;
    d[q] = hi[q] - lo[q];
    if (d[q] > dm) dm = d[q];
  }
  for (int q = 0; q < 3; q++) pmin[q] = lo[q] - 0.5 * (dm - d[q]);
  *D = dm;
}

void fmb_tree_destroy(fmb_tree *t) {
  if (!t) return;
  for (int l = 0; l < FMB_MAXLEV; l++) {
    fmbh_free(t->d_keys[l]);
    fmbh_free(t->d_parent[l]);
    fmbh_free(t->d_cstart[l]);
    fmbh_free(t->d_pair_tgt[l]);
    fmbh_free(t->d_pair_src[l]);
    free(t->h_off_start[l]);
  }
  fmbh_free(t->d_perm);
  fmbh_free(t->d_iperm);
  fmbh_free(t->d_box_start);
  fmbh_free(t->d_elem_box);
  fmbh_free(t->d_tri);
  fmbh_free(t->d_geo);
  fmbh_free(t->d_fv);
  fmbh_free(t->d_vnorm);
  fmbh_free(t->d_near_start);
  fmbh_free(t->d_near_box);
  free(t);
}

int fmb_tree_build(const fmb_mesh *m, int s, int depth, void *stream,
                   fmb_tree **out) {
  if (!m || !out || m->nf < 1)
    return fmb_set_error(FMB_ERR_ARG, "fmb_tree_build: invalid arguments");
  if (depth < 0 || depth > 10 || (depth == 0 && s < 1))
    return fmb_set_error(FMB_ERR_ARG, "fmb_tree_build: depth %d s %d", depth, s);
  int rc = fmbh_init();
  if (rc) return rc;
  fmb_tree *t = (fmb_tree *)calloc(1, sizeof(fmb_tree));
  t->nelem = m->nf;
  t->lmin = 2;
  t->stream = stream;
  int n = m->nf;
  double *d_V = NULL, *d_ctr = NULL;
  int *d_F = NULL, *d_key = NULL;
  if ((rc = fmbh_malloc((void **)&d_V, sizeof(double) * 3 * m->nv))) goto fail;
  if ((rc = fmbh_malloc((void **)&d_F, sizeof(int) * 3 * n))) goto fail;
  if ((rc = fmbh_malloc((void **)&d_ctr, sizeof(double) * 3 * n))) goto fail;
  if ((rc = fmbh_malloc((void **)&d_key, sizeof(int) * n))) goto fail;
  if ((rc = fmbh_malloc((void **)&t->d_perm, sizeof(int) * n))) goto fail;
  if ((rc = fmbh_h2d(d_V, m->V, sizeof(double) * 3 * m->nv, stream))) goto fail;
  if ((rc = fmbh_h2d(d_F, m->F, sizeof(int) * 3 * n, stream))) goto fail;
  {
    double lo[3], hi[3];
    if ((rc = fmbh_centroid_bbox(d_V, d_F, n, d_ctr, lo, hi, stream))) goto fail;
    bounding_cube(lo, hi, t->pmin, &t->D);
  }
  if ((rc = fmbh_morton_sort(d_ctr, n, t->pmin, t->D, d_key, t->d_perm, stream)))
    goto fail;
  if (depth > 0) {
    t->lmax = depth;
  } else {
    int L = 0;
    if ((rc = fmbh_deter_lmax(d_key, n, s, &L, stream))) goto fail;
    t->lmax = L;
  }
  if (t->lmax < 2) t->lmax = 2; /* lmin = 2 (translation.cu:2976) */
  if ((rc = fmbh_leaf_level(t, d_key, stream))) goto fail;
  for (int l = t->lmax - 1; l >= 2; l--)
    if ((rc = fmbh_parent_level(t, l, stream))) goto fail;
  for (int l = 2; l <= t->lmax; l++)
    if ((rc = fmbh_m2l_pairs(t, l, stream))) goto fail;
  if ((rc = fmbh_near_lists(t, stream))) goto fail;
  if ((rc = fmbh_sorted_geometry(t, d_V, d_F, stream))) goto fail;
  /* vertex normals: sums of the faces' (n0-n2)x(n1-n2) (area weights) */
  {
    double *vn = (double *)calloc((size_t)3 * m->nv, sizeof(double));
    for (int f = 0; f < n; f++) {
      const int *id = m->F + 3 * f;
      const double *P0 = m->V + 3 * id[0], *P1 = m->V + 3 * id[1], *P2 = m->V + 3 * id[2];
      double a[3], b[3];
      for (int q = 0; q < 3; q++) {
        a[q] = P0[q] - P2[q];
        b[q] = P1[q] - P2[q];
      }
      double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                     a[0] * b[1] - a[1] * b[0]};
      for (int v = 0; v < 3; v++)
        for (int q = 0; q < 3; q++) vn[3 * id[v] + q] += c[q];
    }
    for (int v = 0; v < m->nv; v++) {
      double s2 = sqrt(vn[3 * v] * vn[3 * v] + vn[3 * v + 1] * vn[3 * v + 1] +
                       vn[3 * v + 2] * vn[3 * v + 2]);
      if (s2 > 0)
        for (int q = 0; q < 3; q++) vn[3 * v + q] /= s2;
    }
    t->nvert = m->nv;
    rc = fmbh_malloc((void **)&t->d_vnorm, sizeof(double) * 3 * m->nv);
    if (!rc) rc = fmbh_h2d(t->d_vnorm, vn, sizeof(double) * 3 * m->nv, stream);
    if (!rc) rc = fmbh_sync(stream);
    free(vn);
    if (rc) goto fail;
  }
  {
    int L = t->lmax, nb = t->nbox[L];
    int *bs = (int *)malloc(sizeof(int) * (nb + 1));
    if ((rc = fmbh_d2h(bs, t->d_box_start, sizeof(int) * (nb + 1), stream))) {
      free(bs);
      goto fail;
    }
    int mx = 0;
    for (int b = 0; b < nb; b++)
      if (bs[b + 1] - bs[b] > mx) mx = bs[b + 1] - bs[b];
    t->max_leaf_elems = mx;
    free(bs);
  }
  fmbh_free(d_V);
  fmbh_free(d_F);
  fmbh_free(d_ctr);
  fmbh_free(d_key);
  *out = t;
  return FMB_OK;
fail:
  fmbh_free(d_V);
  fmbh_free(d_F);
  fmbh_free(d_ctr);
  fmbh_free(d_key);
  fmb_tree_destroy(t);
  return rc;
}

int fmb_tree_info(const fmb_tree *t, int *lmin, int *lmax, double *D,
                  double *pmin) {
  if (!t) return fmb_set_error(FMB_ERR_ARG, "null tree");
  if (lmin) *lmin = t->lmin;
  if (lmax) *lmax = t->lmax;
  if (D) *D = t->D;
  if (pmin) memcpy(pmin, t->pmin, sizeof(double) * 3);
  return FMB_OK;
}

int fmb_tree_level(const fmb_tree *t, int l, int *nbox, int *keys) {
  if (!t || l < t->lmin || l > t->lmax)
    return fmb_set_error(FMB_ERR_ARG, "level %d outside the tree", l);
  if (nbox) *nbox = t->nbox[l];
  if (keys) return fmbh_d2h(keys, t->d_keys[l], sizeof(int) * t->nbox[l], t->stream);
  return FMB_OK;
}

int fmb_tree_pairs(const fmb_tree *t, int l, int *npairs, int *tgt, int *src,
                   int *off) {
  if (!t || l < t->lmin || l > t->lmax)
    return fmb_set_error(FMB_ERR_ARG, "level %d outside the tree", l);
  int np = t->npairs[l], rc;
  if (npairs) *npairs = np;
  if (tgt && (rc = fmbh_d2h(tgt, t->d_pair_tgt[l], sizeof(int) * np, t->stream)))
    return rc;
  if (src && (rc = fmbh_d2h(src, t->d_pair_src[l], sizeof(int) * np, t->stream)))
    return rc;
  if (off && np > 0)
    for (int o = 0; o < FMB_NOFF; o++)
      for (int q = t->h_off_start[l][o]; q < t->h_off_start[l][o + 1]; q++)
        off[q] = o;
  return FMB_OK;
}

int fmb_tree_num_sr(const fmb_tree *t, int l, int *num) {
  if (!t || l < t->lmin || l > t->lmax || !num)
    return fmb_set_error(FMB_ERR_ARG, "level %d outside the tree", l);
  int np = t->npairs[l];
  int *tg = (int *)malloc(sizeof(int) * (np > 0 ? np : 1));
  int rc = fmbh_d2h(tg, t->d_pair_tgt[l], sizeof(int) * np, t->stream);
  if (!rc) {
    memset(num, 0, sizeof(int) * t->nbox[l]);
    for (int q = 0; q < np; q++) num[tg[q]]++;
  }
  free(tg);
  return rc;
}

int fmb_tree_elem_box(const fmb_tree *t, int *box, int *perm) {
  if (!t) return fmb_set_error(FMB_ERR_ARG, "null tree");
  int n = t->nelem, rc;
  int *p = (int *)malloc(sizeof(int) * n), *eb = (int *)malloc(sizeof(int) * n);
  if ((rc = fmbh_d2h(p, t->d_perm, sizeof(int) * n, t->stream))) goto done;
  if ((rc = fmbh_d2h(eb, t->d_elem_box, sizeof(int) * n, t->stream))) goto done;
  if (perm) memcpy(perm, p, sizeof(int) * n);
  if (box)
    for (int i = 0; i < n; i++) box[p[i]] = eb[i];
done:
  free(p);
  free(eb);
  return rc;
}

int fmb_tree_vertex_normals(const fmb_tree *t, double *normals) {
  if (!t || !normals) return fmb_set_error(FMB_ERR_ARG, "fmb_tree_vertex_normals");
  return fmbh_d2h(normals, t->d_vnorm, sizeof(double) * 3 * t->nvert, t->stream);
}
