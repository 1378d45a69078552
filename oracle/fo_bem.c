/*
 * fo_bem.c -- Burton-Miller collocation BEM, FMM matvec, GMRES and the Mie
 * series of the oracle (FP64, OpenMP).
 *
 * TEST INFRASTRUCTURE ONLY (see fmb_oracle.h).
 *
 * Discretisation (integral.cu:80-751, structs.h:29-35): flat triangles,
 * one constant unknown per triangle, collocation at the centroid.  With
 * dp/dn = sigma p (sigma = 0: rigid) the Burton-Miller row i is
 *   (1/2 + alpha sigma_i/2) u_i
 *     + sum_j [ -K_ij - alpha T_ij + sigma_j (S_ij + alpha K'_ij) ] u_j
 *   = p_inc(x_i) + alpha dp_inc/dn(x_i)
 * (S: int G, K: int dG/dn_y, K': int dG/dn_x, T: f.p. int d2G/dn_x dn_y).
 *
 * FMM (translation.cu:3412-3769, completed): P2M as
 *   ik sum_q w_q [-dR_n^{-m}/dn_y + sigma R_n^{-m}](y_q - c)  per element
 * (bottomLevelCoeff_L / triElemIntegral_pRpn, integral.cu:333-365, 821-836),
 * M2M "SS" with (R|R) coaxial matrices, M2L "SR" over I4, L2L "RR", all by
 * rotate-coaxial-rotate; L2P and the near field (the 27-box neighbourhood of
 * each leaf) complete the operator that the reference never finishes.
 * Per-level orders follow truncNum(k, 0.05, 1.5, side) (translation.cu:3432)
 * unless given.
 *
 * Pinned by tests/test_oracle_bem.py: FMM matvec against the dense O(N^2)
 * matvec with the same quadrature, the dense solve against the rigid-sphere
 * Mie series, GMRES against numpy.linalg.solve on small systems.
 */
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>
#include <string.h>

#include "fmb_oracle.h"

#define MAXQ 4096

struct fo_bem {
  int N;
  double k;
  fo_cplx alpha;
  fo_cplx *sig;
  double *tri, *ctr, *nrm, *area;
  int nq;
  double bary[3 * MAXQ], wq[MAXQ];
  int nqn; /* refined rule for close pairs */
  double baryn[3 * MAXQ], wqn[MAXQ];
  double near_ratio;
  double *rad; /* panel circumradius about the centroid */
  int nedge;
  /* FMM */
  fo_tree *tree;
  int L;
  int p[16];
  fo_cplx *Pb, *Eb; /* [N][pL^2], original element order */
  double *bc[16];   /* physical box centres */
  int *par[16];     /* parent box index (level l-1) */
  int *il_start[16], *il_src[16], *il_off[16];
  int *nb_start, *nb_list; /* leaf neighbourhoods incl. self */
  int *nr_start, *nr_col;  /* near rows (original order) */
  fo_cplx *nr_val;
  /* operators */
  int nbeta;
  double beta_cos[400];
  double *Hb[16]; /* per level: [nbeta][rot_size(p_l')] */
  int hsz[16];
  double *Csr[16]; /* per level: [15 distances][coax_size(p_l)] complex */
  int dist_of_d2[28];
  double *Crr[16]; /* per level l>=3: (R|R) at p_{l-1}, side_l*sqrt(3)/2 */
};

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

int fo_num_threads(void) { return omp_get_max_threads(); }

int fo_bem_create(const double *V, int nv, const int *F, int nf, double k,
                  const double *alpha, const double *sig_re,
                  const double *sig_im, int rule, int order, int nedge,
                  fo_bem **out) {
  (void)nv;
  fo_bem *b = (fo_bem *)calloc(1, sizeof(fo_bem));
  b->N = nf;
  b->k = k;
  b->alpha = alpha[0] + I * alpha[1];
  b->nedge = nedge;
  b->sig = (fo_cplx *)calloc(nf, sizeof(fo_cplx));
  if (sig_re)
    for (int i = 0; i < nf; i++)
      b->sig[i] = sig_re[i] + I * (sig_im ? sig_im[i] : 0.0);
  b->tri = (double *)malloc(sizeof(double) * 9 * nf);
  b->ctr = (double *)malloc(sizeof(double) * 3 * nf);
  b->nrm = (double *)malloc(sizeof(double) * 3 * nf);
  b->area = (double *)malloc(sizeof(double) * nf);
  for (int i = 0; i < nf; i++) {
    double *t = b->tri + 9 * i;
    for (int v = 0; v < 3; v++)
      for (int c = 0; c < 3; c++) t[3 * v + c] = V[3 * F[3 * i + v] + c];
    double a[3], bb[3], cr[3];
    for (int c = 0; c < 3; c++) {
      a[c] = t[c] - t[6 + c];
      bb[c] = t[3 + c] - t[6 + c];
      b->ctr[3 * i + c] = t[c] + 2.0 / 3.0 * (0.5 * (t[3 + c] + t[6 + c]) - t[c]);
    }
    cr[0] = a[1] * bb[2] - a[2] * bb[1];
    cr[1] = a[2] * bb[0] - a[0] * bb[2];
    cr[2] = a[0] * bb[1] - a[1] * bb[0];
    double J = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    for (int c = 0; c < 3; c++) b->nrm[3 * i + c] = cr[c] / J;
    b->area[i] = 0.5 * J;
  }
  if (rule == FO_RULE_COLLAPSED && order * order > MAXQ) order = 64;
  b->nq = fo_tri_rule(rule, order, b->bary, b->wq);
  b->nqn = 0;
  b->near_ratio = 0.0;
  b->rad = (double *)malloc(sizeof(double) * nf);
  for (int i = 0; i < nf; i++) {
    double r = 0.0;
    for (int v = 0; v < 3; v++) {
      double d2 = 0.0;
      for (int c = 0; c < 3; c++) {
        double d = b->tri[9 * i + 3 * v + c] - b->ctr[3 * i + c];
        d2 += d * d;
      }
      if (d2 > r) r = d2;
    }
    b->rad[i] = sqrt(r);
  }
  *out = b;
  return 0;
}

int fo_bem_n(const fo_bem *b) { return b->N; }

/* Close pairs (|x_i - c_j| < ratio * circumradius_j) use a refined rule:
 * the hypersingular kernel on neighbouring panels is not resolved by the
 * 7-point rule. */
void fo_bem_set_near(fo_bem *b, int rule, int order, double ratio) {
  if (rule == FO_RULE_COLLAPSED && order * order > MAXQ) order = 64;
  b->nqn = fo_tri_rule(rule, order, b->baryn, b->wqn);
  b->near_ratio = ratio;
}

void fo_bem_centroids(const fo_bem *b, double *c, double *nrm, double *area) {
  memcpy(c, b->ctr, sizeof(double) * 3 * b->N);
  memcpy(nrm, b->nrm, sizeof(double) * 3 * b->N);
  memcpy(area, b->area, sizeof(double) * b->N);
}

static fo_cplx entry(const fo_bem *b, int i, int j) {
  double re4[4], im4[4];
  fo_cplx v;
  if (i == j) {
    fo_panel_self(b->k, b->tri + 9 * j, b->nedge, re4, im4);
  } else {
    /* regular panel rule, refined for close pairs */
    const double *t = b->tri + 9 * j;
    const double *bary = b->bary, *wq = b->wq;
    int nq = b->nq;
    if (b->nqn > 0) {
      double d2 = 0.0;
      for (int c = 0; c < 3; c++) {
        double d = b->ctr[3 * i + c] - b->ctr[3 * j + c];
        d2 += d * d;
      }
      if (d2 < b->near_ratio * b->near_ratio * b->rad[j] * b->rad[j]) {
        bary = b->baryn;
        wq = b->wqn;
        nq = b->nqn;
      }
    }
    for (int q = 0; q < 4; q++) re4[q] = im4[q] = 0.0;
    for (int q = 0; q < nq; q++) {
      double y[3], kr[4], ki[4];
      for (int c = 0; c < 3; c++)
        y[c] = bary[3 * q] * t[c] + bary[3 * q + 1] * t[3 + c] +
               bary[3 * q + 2] * t[6 + c];
      fo_kernels(b->k, b->ctr + 3 * i, b->nrm + 3 * i, y, b->nrm + 3 * j, kr,
                 ki);
      double w = wq[q] * b->area[j];
      for (int c = 0; c < 4; c++) {
        re4[c] += w * kr[c];
        im4[c] += w * ki[c];
      }
    }
  }
  fo_cplx S = re4[0] + I * im4[0], K = re4[1] + I * im4[1];
  fo_cplx Kp = re4[2] + I * im4[2], T = re4[3] + I * im4[3];
  v = -K - b->alpha * T + b->sig[j] * (S + b->alpha * Kp);
  if (i == j) v += 0.5 + 0.5 * b->alpha * b->sig[i];
  return v;
}

/* The field kernel of element j at point yq: int_j (dG/dn_y - sigma_j G) dS_y
 * with the panel rule of the matvec (refined when |yq - c_j| < ratio *
 * radius_j).  The same kernels and signs as the collocation entries above. */
static fo_cplx field_panel(const fo_bem *b, int j, const double *yq) {
  const double *t = b->tri + 9 * j;
  const double *bary = b->bary, *wq = b->wq;
  int nq = b->nq;
  double d2 = 0.0, zero[3] = {0, 0, 0};
  for (int c = 0; c < 3; c++) {
    double d = yq[c] - b->ctr[3 * j + c];
    d2 += d * d;
  }
  if (b->nqn > 0 && d2 < b->near_ratio * b->near_ratio * b->rad[j] * b->rad[j]) {
    bary = b->baryn;
    wq = b->wqn;
    nq = b->nqn;
  }
  fo_cplx S = 0.0, K = 0.0;
  for (int g = 0; g < nq; g++) {
    double y[3], kr[4], ki[4];
    for (int c = 0; c < 3; c++)
      y[c] = bary[3 * g] * t[c] + bary[3 * g + 1] * t[3 + c] + bary[3 * g + 2] * t[6 + c];
    fo_kernels(b->k, yq, zero, y, b->nrm + 3 * j, kr, ki);
    double w = wq[g] * b->area[j];
    S += w * (kr[0] + I * ki[0]);
    K += w * (kr[1] + I * ki[1]);
  }
  return K - b->sig[j] * S;
}

static void field_store(const fo_bem *b, int nrhs, const fo_cplx *acc,
                        const double *yq, int q, int npts, const double *dirs,
                        double *ore, double *oim) {
  for (int r = 0; r < nrhs; r++) {
    fo_cplx v = acc[r];
    if (dirs) {
      const double *d = dirs + 3 * r;
      v += cexp(I * b->k * (d[0] * yq[0] + d[1] * yq[1] + d[2] * yq[2]));
    }
    ore[(size_t)r * npts + q] = creal(v);
    oim[(size_t)r * npts + q] = cimag(v);
  }
}

/* Pressure at field points off the surface by the exterior representation
 * formula p(y) = p_inc(y) + sum_j x_j int_j (dG/dn_y - sigma_j G) dS_y, the
 * panel rule of the matvec (refined when |y - c_j| < ratio * radius_j). */
void fo_bem_field(const fo_bem *b, int nrhs, const double *xre, const double *xim,
                  const double *pts, int npts, const double *dirs, double *ore,
                  double *oim) {
  int N = b->N;
#pragma omp parallel for schedule(dynamic, 1)
  for (int q = 0; q < npts; q++) {
    const double *yq = pts + 3 * q;
    fo_cplx *acc = (fo_cplx *)calloc(nrhs, sizeof(fo_cplx));
    for (int j = 0; j < N; j++) {
      fo_cplx gj = field_panel(b, j, yq);
      for (int r = 0; r < nrhs; r++)
        acc[r] += gj * (xre[(size_t)r * N + j] + I * xim[(size_t)r * N + j]);
    }
    field_store(b, nrhs, acc, yq, q, npts, dirs, ore, oim);
    free(acc);
  }
}

void fo_bem_dense_row(const fo_bem *b, int i, double *re, double *im) {
  for (int j = 0; j < b->N; j++) {
    fo_cplx v = entry(b, i, j);
    re[j] = creal(v);
    im[j] = cimag(v);
  }
}

void fo_bem_dense_matvec(const fo_bem *b, int nrhs, const double *xre,
                         const double *xim, double *yre, double *yim) {
  int N = b->N;
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < N; i++) {
    fo_cplx *acc = (fo_cplx *)calloc(nrhs, sizeof(fo_cplx));
    for (int j = 0; j < N; j++) {
      fo_cplx a = entry(b, i, j);
      for (int r = 0; r < nrhs; r++)
        acc[r] += a * (xre[(size_t)r * N + j] + I * xim[(size_t)r * N + j]);
    }
    for (int r = 0; r < nrhs; r++) {
      yre[(size_t)r * N + i] = creal(acc[r]);
      yim[(size_t)r * N + i] = cimag(acc[r]);
    }
    free(acc);
  }
}

void fo_bem_rhs_plane(const fo_bem *b, const double *dir, double *re,
                      double *im) {
  for (int i = 0; i < b->N; i++) {
    const double *x = b->ctr + 3 * i, *n = b->nrm + 3 * i;
    double dx = dir[0] * x[0] + dir[1] * x[1] + dir[2] * x[2];
    double dn = dir[0] * n[0] + dir[1] * n[1] + dir[2] * n[2];
    fo_cplx e = cexp(I * b->k * dx);
    fo_cplx v = e * (1.0 + b->alpha * I * b->k * dn);
    re[i] = creal(v);
    im[i] = cimag(v);
  }
}

/* ------------------------------------------------------------------ FMM */

static int beta_index(fo_bem *b, double cb) {
  for (int i = 0; i < b->nbeta; i++)
    if (fabs(b->beta_cos[i] - cb) < 1e-12) return i;
  b->beta_cos[b->nbeta] = cb;
  return b->nbeta++;
}

/* rotate-coaxial-rotate with H(beta), phase phi, coaxial matrix at order P */
static void rcr(int P, const double *H, double phi, const double *Cc, int pin,
                int pout, const fo_cplx *x, fo_cplx *y) {
  int Pm = pin > pout ? pin : pout;
  fo_cplx *a = (fo_cplx *)malloc(sizeof(fo_cplx) * Pm * Pm);
  fo_cplx *c = (fo_cplx *)malloc(sizeof(fo_cplx) * Pm * Pm);
  const fo_cplx *C = (const fo_cplx *)Cc;
  /* forward rotation (alpha, beta, gamma) = (phi, theta, 0) */
  for (int n = 0; n < pin; n++) {
    int off = n * (4 * n * n - 1) / 3, ld = 2 * n + 1;
    for (int mp = -n; mp <= n; mp++) {
      fo_cplx s = 0.0;
      for (int m = -n; m <= n; m++)
        s += H[off + (m + n) * ld + (mp + n)] * cexp(I * m * phi) *
             x[n * n + n + m];
      a[n * n + n + mp] = s;
    }
  }
  /* coaxial */
  int pm = pin < pout ? pin : pout;
  for (int i = 0; i < pout * pout; i++) c[i] = 0.0;
  for (int m = -(pm - 1); m <= pm - 1; m++) {
    int am = m < 0 ? -m : m;
    int off = fo_coax_block(P, am), ld = P - am;
    for (int np = am; np < pout; np++) {
      fo_cplx s = 0.0;
      for (int n = am; n < pin; n++)
        s += C[off + (n - am) * ld + (np - am)] * a[n * n + n + m];
      c[np * np + np + m] = s;
    }
  }
  /* back rotation (alpha, beta, gamma) = (0, theta, phi) */
  for (int n = 0; n < pout; n++) {
    int off = n * (4 * n * n - 1) / 3, ld = 2 * n + 1;
    for (int mp = -n; mp <= n; mp++) {
      fo_cplx s = 0.0;
      for (int m = -n; m <= n; m++)
        s += H[off + (m + n) * ld + (mp + n)] * c[n * n + n + m];
      y[n * n + n + mp] += cexp(-I * mp * phi) * s;
    }
  }
  free(a);
  free(c);
}

static int dist_class(int d2) {
  static const int tab[15] = {4, 5, 6, 8, 9, 10, 11, 12, 13, 14, 17, 18, 19, 22, 27};
  for (int i = 0; i < 15; i++)
    if (tab[i] == d2) return i;
  return -1;
}

int fo_bem_fmm_setup(fo_bem *b, int s, int depth, const int *plev) {
  int N = b->N;
  /* the tree is built from vertices; rebuild a vertex array from panels */
  double *V = (double *)malloc(sizeof(double) * 9 * N);
  int *F = (int *)malloc(sizeof(int) * 3 * N);
  memcpy(V, b->tri, sizeof(double) * 9 * N);
  for (int i = 0; i < 3 * N; i++) F[i] = i;
  int rc = fo_tree_build(V, 3 * N, F, N, s, depth, &b->tree);
  free(V);
  free(F);
  if (rc) return rc;
  fo_tree *t = b->tree;
  int L = t->lmax;
  b->L = L;
  for (int l = 2; l <= L; l++) {
    double side = t->D * ldexp(1.0, -l);
    b->p[l] = (plev && plev[l] > 0) ? plev[l]
                                    : fo_trunc_num(b->k, 0.05, 1.5, side);
  }
  /* box centres and parents */
  for (int l = 2; l <= L; l++) {
    int nb = t->nbox[l];
    b->bc[l] = (double *)malloc(sizeof(double) * 3 * nb);
    b->par[l] = (int *)malloc(sizeof(int) * nb);
    for (int i = 0; i < nb; i++) {
      double c[3];
      fo_box_center(t->keys[l][i], l, c);
      for (int q = 0; q < 3; q++) b->bc[l][3 * i + q] = t->D * c[q] + t->pmin[q];
      if (l > 2) {
        int pk = t->keys[l][i] >> 3;
        int *hit = (int *)bsearch(&pk, t->keys[l - 1], t->nbox[l - 1],
                                  sizeof(int), cmp_int);
        b->par[l][i] = (int)(hit - t->keys[l - 1]);
      } else {
        b->par[l][i] = -1;
      }
    }
  }
  /* interaction lists */
  for (int l = 2; l <= L; l++) {
    int nb = t->nbox[l];
    b->il_start[l] = (int *)malloc(sizeof(int) * (nb + 1));
    int cap = nb * 189 + 1;
    b->il_src[l] = (int *)malloc(sizeof(int) * cap);
    b->il_off[l] = (int *)malloc(sizeof(int) * cap);
    int cnt = 0, set[216];
    for (int i = 0; i < nb; i++) {
      b->il_start[l][i] = cnt;
      int n = fo_I4(t->keys[l][i], l, set);
      int tx, ty, tz;
      {
        double c[3];
        fo_box_center(t->keys[l][i], l, c);
        tx = (int)floor(c[0] * (1 << l));
        ty = (int)floor(c[1] * (1 << l));
        tz = (int)floor(c[2] * (1 << l));
      }
      for (int j = 0; j < n; j++) {
        int *hit = (int *)bsearch(&set[j], t->keys[l], nb, sizeof(int), cmp_int);
        if (!hit) continue;
        double c[3];
        fo_box_center(set[j], l, c);
        int sx = (int)floor(c[0] * (1 << l)), sy = (int)floor(c[1] * (1 << l)),
            sz = (int)floor(c[2] * (1 << l));
        int dx = tx - sx, dy = ty - sy, dz = tz - sz;
        b->il_src[l][cnt] = (int)(hit - t->keys[l]);
        b->il_off[l][cnt] = (dx + 3) * 49 + (dy + 3) * 7 + (dz + 3);
        cnt++;
      }
    }
    b->il_start[l][nb] = cnt;
  }
  /* leaf neighbourhoods */
  {
    int nb = t->nbox[L];
    b->nb_start = (int *)malloc(sizeof(int) * (nb + 1));
    b->nb_list = (int *)malloc(sizeof(int) * 27 * nb);
    int cnt = 0, nbr[27];
    for (int i = 0; i < nb; i++) {
      b->nb_start[i] = cnt;
      int n = fo_neighbors(t->keys[L][i], L, nbr);
      nbr[n++] = t->keys[L][i];
      qsort(nbr, n, sizeof(int), cmp_int);
      for (int j = 0; j < n; j++) {
        int *hit = (int *)bsearch(&nbr[j], t->keys[L], nb, sizeof(int), cmp_int);
        if (hit) b->nb_list[cnt++] = (int)(hit - t->keys[L]);
      }
    }
    b->nb_start[nb] = cnt;
  }
  /* near rows */
  {
    b->nr_start = (int *)malloc(sizeof(int) * (N + 1));
    long tot = 0;
    for (int i = 0; i < N; i++) {
      b->nr_start[i] = (int)tot;
      int bx = t->elem_box[i];
      for (int q = b->nb_start[bx]; q < b->nb_start[bx + 1]; q++) {
        int nbx = b->nb_list[q];
        tot += t->box_end[nbx] - t->box_start[nbx];
      }
    }
    b->nr_start[N] = (int)tot;
    b->nr_col = (int *)malloc(sizeof(int) * tot);
    b->nr_val = (fo_cplx *)malloc(sizeof(fo_cplx) * tot);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < N; i++) {
      int bx = t->elem_box[i], c = b->nr_start[i];
      for (int q = b->nb_start[bx]; q < b->nb_start[bx + 1]; q++) {
        int nbx = b->nb_list[q];
        for (int e = t->box_start[nbx]; e < t->box_end[nbx]; e++) {
          int j = t->perm[e];
          b->nr_col[c] = j;
          b->nr_val[c] = entry(b, i, j);
          c++;
        }
      }
    }
  }
  /* P2M / L2P bases at the leaf order */
  {
    int pL = b->p[L], ps = pL * pL;
    b->Pb = (fo_cplx *)malloc(sizeof(fo_cplx) * (size_t)N * ps);
    b->Eb = (fo_cplx *)malloc(sizeof(fo_cplx) * (size_t)N * ps);
    fo_cplx ik = I * b->k;
#pragma omp parallel
    {
      double *rr = (double *)malloc(sizeof(double) * ps);
      double *ri = (double *)malloc(sizeof(double) * ps);
      double *dr = (double *)malloc(sizeof(double) * ps);
      double *di = (double *)malloc(sizeof(double) * ps);
#pragma omp for schedule(dynamic, 16)
      for (int j = 0; j < N; j++) {
        const double *c = b->bc[L] + 3 * t->elem_box[j];
        fo_cplx *P = b->Pb + (size_t)j * ps;
        for (int i = 0; i < ps; i++) P[i] = 0.0;
        const double *tr = b->tri + 9 * j;
        for (int q = 0; q < b->nq; q++) {
          double v[3];
          for (int a = 0; a < 3; a++)
            v[a] = b->bary[3 * q] * tr[a] + b->bary[3 * q + 1] * tr[3 + a] +
                   b->bary[3 * q + 2] * tr[6 + a] - c[a];
          fo_rbasis(pL - 1, b->k, v, rr, ri);
          fo_rbasis_dn(pL - 1, b->k, v, b->nrm + 3 * j, dr, di);
          double w = b->wq[q] * b->area[j];
          for (int n = 0; n < pL; n++)
            for (int m = -n; m <= n; m++) {
              int src = n * n + n - m; /* R_n^{-m} */
              fo_cplx R = rr[src] + I * ri[src];
              fo_cplx dR = dr[src] + I * di[src];
              P[n * n + n + m] += ik * w * (-dR + b->sig[j] * R);
            }
        }
        /* L2P: R + alpha dR/dn at the centroid */
        double v[3];
        for (int a = 0; a < 3; a++) v[a] = b->ctr[3 * j + a] - c[a];
        fo_rbasis(pL - 1, b->k, v, rr, ri);
        fo_rbasis_dn(pL - 1, b->k, v, b->nrm + 3 * j, dr, di);
        fo_cplx *E = b->Eb + (size_t)j * ps;
        for (int i = 0; i < ps; i++)
          E[i] = (rr[i] + I * ri[i]) + b->alpha * (dr[i] + I * di[i]);
      }
      free(rr);
      free(ri);
      free(dr);
      free(di);
    }
  }
  /* operators: H per unique cos(beta) at the largest order of each level */
  b->nbeta = 0;
  for (int dx = -3; dx <= 3; dx++)
    for (int dy = -3; dy <= 3; dy++)
      for (int dz = -3; dz <= 3; dz++) {
        int d2 = dx * dx + dy * dy + dz * dz;
        if (!d2) continue;
        beta_index(b, dz / sqrt((double)d2));
      }
  for (int l = 2; l <= L; l++) {
    int P = b->p[l];
    if (l > 2 && b->p[l - 1] > P) P = b->p[l - 1];
    if (l < L && b->p[l + 1] > P) P = b->p[l + 1];
    int rs = fo_rot_size(P);
    b->hsz[l] = rs;
    b->Hb[l] = (double *)malloc(sizeof(double) * rs * b->nbeta);
#pragma omp parallel for
    for (int i = 0; i < b->nbeta; i++)
      fo_rot_H(P, acos(b->beta_cos[i]), b->Hb[l] + (size_t)rs * i);
    double side = t->D * ldexp(1.0, -l);
    int pl = b->p[l], cs = fo_coax_size(pl);
    b->Csr[l] = (double *)malloc(sizeof(double) * 2 * cs * 15);
    static const int tab[15] = {4, 5, 6, 8, 9, 10, 11, 12, 13, 14, 17, 18, 19, 22, 27};
#pragma omp parallel for
    for (int i = 0; i < 15; i++) {
      double *re = (double *)malloc(sizeof(double) * cs);
      double *im = (double *)malloc(sizeof(double) * cs);
      fo_coax_sparse(FO_COAX_SR, pl, b->k, side * sqrt((double)tab[i]), re, im);
      fo_cplx *dst = (fo_cplx *)(b->Csr[l] + 2 * (size_t)cs * i);
      for (int q = 0; q < cs; q++) dst[q] = re[q] + I * im[q];
      free(re);
      free(im);
    }
    if (l > 2) {
      int pp = b->p[l - 1] > b->p[l] ? b->p[l - 1] : b->p[l];
      int c2 = fo_coax_size(pp);
      double *re = (double *)malloc(sizeof(double) * c2);
      double *im = (double *)malloc(sizeof(double) * c2);
      fo_coax_sparse(FO_COAX_RR, pp, b->k, side * sqrt(3.0) / 2.0, re, im);
      b->Crr[l] = (double *)malloc(sizeof(double) * 2 * c2);
      fo_cplx *dst = (fo_cplx *)b->Crr[l];
      for (int q = 0; q < c2; q++) dst[q] = re[q] + I * im[q];
      free(re);
      free(im);
    }
  }
  return 0;
}

int fo_bem_fmm_orders(const fo_bem *b, int *p) {
  for (int l = 0; l < 16; l++) p[l] = (l >= 2 && l <= b->L) ? b->p[l] : 0;
  return b->L;
}

static const double *Hof(const fo_bem *b, int l, double cb) {
  for (int i = 0; i < b->nbeta; i++)
    if (fabs(b->beta_cos[i] - cb) < 1e-12) return b->Hb[l] + (size_t)b->hsz[l] * i;
  return NULL;
}

/* The upward pass of one vector: P2M into the leaves, M2M up to level 2.
 * M[l], l = 2..L, zeroed [nbox_l][p_l^2] arrays. */
static void fmm_up(const fo_bem *b, const fo_cplx *x, fo_cplx **M) {
  const fo_tree *t = b->tree;
  int L = b->L;
  int pL = b->p[L], ps = pL * pL;
  /* P2M */
#pragma omp parallel for schedule(dynamic, 8)
  for (int bx = 0; bx < t->nbox[L]; bx++) {
    fo_cplx *dst = M[L] + (size_t)bx * ps;
    for (int e = t->box_start[bx]; e < t->box_end[bx]; e++) {
      int j = t->perm[e];
      const fo_cplx *P = b->Pb + (size_t)j * ps;
      for (int i = 0; i < ps; i++) dst[i] += P[i] * x[j];
    }
  }
  /* M2M: child at l -> parent at l-1, t = parent - child */
  for (int l = L; l > 2; l--) {
    int pc = b->p[l], pp = b->p[l - 1];
    int P = pc > pp ? pc : pp;
    double side = t->D * ldexp(1.0, -l);
    int np = t->nbox[l - 1];
    int *cst = (int *)malloc(sizeof(int) * (np + 1));
    for (int q = 0, cb = 0; q <= np; q++) {
      while (cb < t->nbox[l] && b->par[l][cb] < q) cb++;
      cst[q] = cb;
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (int pb = 0; pb < np; pb++) {
      for (int cb = cst[pb]; cb < cst[pb + 1]; cb++) {
        int oct = t->keys[l][cb] & 7;
        double d[3] = {-(((oct >> 2) & 1) - 0.5), -(((oct >> 1) & 1) - 0.5),
                       -((oct & 1) - 0.5)};
        double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const double *H = Hof(b, l, d[2] / r);
        (void)side;
        rcr(P, H, atan2(d[1], d[0]), b->Crr[l], pc, pp,
            M[l] + (size_t)cb * pc * pc, M[l - 1] + (size_t)pb * pp * pp);
      }
    }
    free(cst);
  }
}

/* One FMM matvec: P2M, M2M, M2L + L2L, L2P + near field. */
static void fmm_one(const fo_bem *b, const fo_cplx *x, fo_cplx *y) {
  const fo_tree *t = b->tree;
  int L = b->L, N = b->N;
  fo_cplx *M[16], *Lc[16];
  for (int l = 2; l <= L; l++) {
    size_t n = (size_t)t->nbox[l] * b->p[l] * b->p[l];
    M[l] = (fo_cplx *)calloc(n, sizeof(fo_cplx));
    Lc[l] = (fo_cplx *)calloc(n, sizeof(fo_cplx));
  }
  int pL = b->p[L], ps = pL * pL;
  fmm_up(b, x, M);
  /* M2L and L2L */
  for (int l = 2; l <= L; l++) {
    int pl = b->p[l], cs = fo_coax_size(pl);
    int nb = t->nbox[l];
#pragma omp parallel for schedule(dynamic, 4)
    for (int bx = 0; bx < nb; bx++) {
      fo_cplx *dst = Lc[l] + (size_t)bx * pl * pl;
      for (int q = b->il_start[l][bx]; q < b->il_start[l][bx + 1]; q++) {
        int s = b->il_src[l][q], o = b->il_off[l][q];
        int dx = o / 49 - 3, dy = (o / 7) % 7 - 3, dz = o % 7 - 3;
        int d2 = dx * dx + dy * dy + dz * dz;
        const double *H = Hof(b, l, dz / sqrt((double)d2));
        const double *C = b->Csr[l] + 2 * (size_t)cs * dist_class(d2);
        rcr(pl, H, atan2((double)dy, (double)dx), C, pl, pl,
            M[l] + (size_t)s * pl * pl, dst);
      }
      if (l > 2) {
        int pp = b->p[l - 1];
        int P = pp > pl ? pp : pl;
        int oct = t->keys[l][bx] & 7;
        double d[3] = {((oct >> 2) & 1) - 0.5, ((oct >> 1) & 1) - 0.5,
                       (oct & 1) - 0.5};
        double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const double *H = Hof(b, l, d[2] / r);
        rcr(P, H, atan2(d[1], d[0]), b->Crr[l], pp, pl,
            Lc[l - 1] + (size_t)b->par[l][bx] * pp * pp, dst);
      }
    }
  }
  /* L2P + near */
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < N; i++) {
    int bx = t->elem_box[i];
    const fo_cplx *E = b->Eb + (size_t)i * ps;
    const fo_cplx *D = Lc[L] + (size_t)bx * ps;
    fo_cplx s = 0.0;
    for (int q = 0; q < ps; q++) s += E[q] * D[q];
    for (int q = b->nr_start[i]; q < b->nr_start[i + 1]; q++)
      s += b->nr_val[q] * x[b->nr_col[q]];
    y[i] = s;
  }
  for (int l = 2; l <= L; l++) {
    free(M[l]);
    free(Lc[l]);
  }
}

void fo_bem_fmm_matvec(const fo_bem *b, int nrhs, const double *xre,
                       const double *xim, double *yre, double *yim) {
  int N = b->N;
  fo_cplx *x = (fo_cplx *)malloc(sizeof(fo_cplx) * N);
  fo_cplx *y = (fo_cplx *)malloc(sizeof(fo_cplx) * N);
  for (int r = 0; r < nrhs; r++) {
    for (int i = 0; i < N; i++)
      x[i] = xre[(size_t)r * N + i] + I * xim[(size_t)r * N + i];
    fmm_one(b, x, y);
    for (int i = 0; i < N; i++) {
      yre[(size_t)r * N + i] = creal(y[i]);
      yim[(size_t)r * N + i] = cimag(y[i]);
    }
  }
  free(x);
  free(y);
}

/* Pressure at field points through the tree (a treecode over the FMM's
 * multipole expansions).  Per point y, from level 2 down: a box whose
 * index differs by 2 or more in some coordinate from the index of the
 * level-l box that contains y (the separation of the M2L lists) adds its
 * expansion; a box next to it passes its children to the next level; the
 * leaf boxes next to y's leaf box add their panels by field_panel.  The
 * P2M basis expands the matrix kernel -dG/dn_y + sigma G, the negative of
 * the field kernel, so a box adds -sum_{n<p_l} M_n^m S_n^m(y - c_B) with
 * S_n^m = h_n(kr) Y_n^m.  Points outside the bounding cube have indices
 * outside [0, 2^l) and are handled alike. */
void fo_bem_field_fmm(const fo_bem *b, int nrhs, const double *xre,
                      const double *xim, const double *pts, int npts,
                      const double *dirs, double *ore, double *oim) {
  const fo_tree *t = b->tree;
  int L = b->L, N = b->N;
  int *cst[16] = {0};
  for (int l = 2; l < L; l++) {
    int nb = t->nbox[l];
    cst[l] = (int *)malloc(sizeof(int) * (nb + 1));
    for (int q = 0, cb = 0; q <= nb; q++) {
      while (cb < t->nbox[l + 1] && b->par[l + 1][cb] < q) cb++;
      cst[l][q] = cb;
    }
  }
  /* M[l][box][r][p_l^2] from one upward pass per vector */
  fo_cplx *M[16] = {0}, *Mr[16] = {0};
  for (int l = 2; l <= L; l++) {
    size_t ps = (size_t)b->p[l] * b->p[l];
    M[l] = (fo_cplx *)malloc(sizeof(fo_cplx) * (size_t)t->nbox[l] * nrhs * ps);
    Mr[l] = (fo_cplx *)malloc(sizeof(fo_cplx) * (size_t)t->nbox[l] * ps);
  }
  fo_cplx *x = (fo_cplx *)malloc(sizeof(fo_cplx) * N);
  for (int r = 0; r < nrhs; r++) {
    for (int i = 0; i < N; i++)
      x[i] = xre[(size_t)r * N + i] + I * xim[(size_t)r * N + i];
    for (int l = 2; l <= L; l++)
      memset(Mr[l], 0, sizeof(fo_cplx) * (size_t)t->nbox[l] * b->p[l] * b->p[l]);
    fmm_up(b, x, Mr);
    for (int l = 2; l <= L; l++) {
      size_t ps = (size_t)b->p[l] * b->p[l];
      for (int bx = 0; bx < t->nbox[l]; bx++)
        memcpy(M[l] + ((size_t)bx * nrhs + r) * ps, Mr[l] + bx * ps,
               sizeof(fo_cplx) * ps);
    }
  }
  int pmax = 0;
  for (int l = 2; l <= L; l++) pmax = b->p[l] > pmax ? b->p[l] : pmax;
  int ncap = t->nbox[2] > 216 ? t->nbox[2] : 216;
#pragma omp parallel for schedule(dynamic, 1)
  for (int q = 0; q < npts; q++) {
    const double *yq = pts + 3 * q;
    fo_cplx *acc = (fo_cplx *)calloc(nrhs, sizeof(fo_cplx));
    double *sre = (double *)malloc(sizeof(double) * pmax * pmax);
    double *sim = (double *)malloc(sizeof(double) * pmax * pmax);
    int *cand = (int *)malloc(sizeof(int) * ncap);
    int *next = (int *)malloc(sizeof(int) * ncap);
    int nc = t->nbox[2];
    for (int i = 0; i < nc; i++) cand[i] = i;
    double u[3];
    for (int c = 0; c < 3; c++) u[c] = (yq[c] - t->pmin[c]) / t->D;
    for (int l = 2; l <= L && nc > 0; l++) {
      int p = b->p[l], nn = 0;
      double sc = ldexp(1.0, l);
      double ty[3] = {floor(u[0] * sc), floor(u[1] * sc), floor(u[2] * sc)};
      for (int i = 0; i < nc; i++) {
        int bx = cand[i];
        double c[3], sep = 0.0;
        fo_box_center(t->keys[l][bx], l, c);
        for (int a = 0; a < 3; a++) {
          double d = fabs(floor(c[a] * sc) - ty[a]);
          sep = d > sep ? d : sep;
        }
        if (sep >= 2.0) {
          double v[3];
          for (int a = 0; a < 3; a++) v[a] = yq[a] - b->bc[l][3 * bx + a];
          fo_sbasis(p - 1, b->k, v, sre, sim);
          for (int r = 0; r < nrhs; r++) {
            const fo_cplx *Mb = M[l] + ((size_t)bx * nrhs + r) * p * p;
            fo_cplx s = 0.0;
            for (int f = 0; f < p * p; f++) s += Mb[f] * (sre[f] + I * sim[f]);
            acc[r] -= s;
          }
        } else if (l < L) {
          for (int cb = cst[l][bx]; cb < cst[l][bx + 1]; cb++) next[nn++] = cb;
        } else {
          for (int e = t->box_start[bx]; e < t->box_end[bx]; e++) {
            int j = t->perm[e];
            fo_cplx gj = field_panel(b, j, yq);
            for (int r = 0; r < nrhs; r++)
              acc[r] += gj * (xre[(size_t)r * N + j] + I * xim[(size_t)r * N + j]);
          }
        }
      }
      for (int i = 0; i < nn; i++) cand[i] = next[i];
      nc = nn;
    }
    field_store(b, nrhs, acc, yq, q, npts, dirs, ore, oim);
    free(acc);
    free(sre);
    free(sim);
    free(cand);
    free(next);
  }
  for (int l = 2; l <= L; l++) {
    free(M[l]);
    free(Mr[l]);
    free(cst[l]);
  }
  free(x);
}

void fo_bem_near_matvec(const fo_bem *b, int nrhs, const double *xre,
                        const double *xim, double *yre, double *yim) {
  int N = b->N;
  for (int r = 0; r < nrhs; r++) {
#pragma omp parallel for
    for (int i = 0; i < N; i++) {
      fo_cplx s = 0.0;
      for (int q = b->nr_start[i]; q < b->nr_start[i + 1]; q++) {
        int j = b->nr_col[q];
        s += b->nr_val[q] * (xre[(size_t)r * N + j] + I * xim[(size_t)r * N + j]);
      }
      yre[(size_t)r * N + i] = creal(s);
      yim[(size_t)r * N + i] = cimag(s);
    }
  }
}

void fo_bem_free(fo_bem *b) {
  if (!b) return;
  free(b->sig);
  free(b->tri);
  free(b->ctr);
  free(b->nrm);
  free(b->area);
  free(b->rad);
  free(b->Pb);
  free(b->Eb);
  for (int l = 0; l < 16; l++) {
    free(b->bc[l]);
    free(b->par[l]);
    free(b->il_start[l]);
    free(b->il_src[l]);
    free(b->il_off[l]);
    free(b->Hb[l]);
    free(b->Csr[l]);
    free(b->Crr[l]);
  }
  free(b->nb_start);
  free(b->nb_list);
  free(b->nr_start);
  free(b->nr_col);
  free(b->nr_val);
  fo_tree_free(b->tree);
  free(b);
}

/* ---------------------------------------------------------------- GMRES */

static void apply_op(const fo_bem *b, int use_fmm, const fo_cplx *x,
                     fo_cplx *y) {
  int N = b->N;
  double *xr = (double *)malloc(sizeof(double) * N);
  double *xi = (double *)malloc(sizeof(double) * N);
  double *yr = (double *)malloc(sizeof(double) * N);
  double *yi = (double *)malloc(sizeof(double) * N);
  for (int i = 0; i < N; i++) {
    xr[i] = creal(x[i]);
    xi[i] = cimag(x[i]);
  }
  if (use_fmm)
    fo_bem_fmm_matvec(b, 1, xr, xi, yr, yi);
  else
    fo_bem_dense_matvec(b, 1, xr, xi, yr, yi);
  for (int i = 0; i < N; i++) y[i] = yr[i] + I * yi[i];
  free(xr);
  free(xi);
  free(yr);
  free(yi);
}

static double vnorm(int N, const fo_cplx *v) {
  double s = 0.0;
  for (int i = 0; i < N; i++) s += creal(v[i] * conj(v[i]));
  return sqrt(s);
}

/* Restarted GMRES(m), Saad & Schultz 1986, modified Gram-Schmidt and
 * complex Givens rotations; convergence on the true residual at restarts. */
int fo_gmres(const fo_bem *b, int use_fmm, int m, double tol, int maxit,
             const double *bre, const double *bim, double *xre, double *xim,
             int *iters, double *relres) {
  int N = b->N;
  fo_cplx *rhs = (fo_cplx *)malloc(sizeof(fo_cplx) * N);
  fo_cplx *x = (fo_cplx *)malloc(sizeof(fo_cplx) * N);
  fo_cplx *r = (fo_cplx *)malloc(sizeof(fo_cplx) * N);
  fo_cplx *w = (fo_cplx *)malloc(sizeof(fo_cplx) * N);
  fo_cplx *Vk = (fo_cplx *)malloc(sizeof(fo_cplx) * (size_t)N * (m + 1));
  fo_cplx *Hh = (fo_cplx *)calloc((size_t)(m + 1) * m, sizeof(fo_cplx));
  fo_cplx *cs = (fo_cplx *)malloc(sizeof(fo_cplx) * m);
  fo_cplx *sn = (fo_cplx *)malloc(sizeof(fo_cplx) * m);
  fo_cplx *g = (fo_cplx *)malloc(sizeof(fo_cplx) * (m + 1));
  for (int i = 0; i < N; i++) {
    rhs[i] = bre[i] + I * bim[i];
    x[i] = xre[i] + I * xim[i];
  }
  double bn = vnorm(N, rhs);
  if (bn == 0.0) bn = 1.0;
  int it = 0, conv = 0;
  double rr = 1.0;
  while (it < maxit) {
    apply_op(b, use_fmm, x, w);
    for (int i = 0; i < N; i++) r[i] = rhs[i] - w[i];
    double beta = vnorm(N, r);
    rr = beta / bn;
    if (rr <= tol) {
      conv = 1;
      break;
    }
    for (int i = 0; i < N; i++) Vk[i] = r[i] / beta;
    for (int i = 0; i <= m; i++) g[i] = 0.0;
    g[0] = beta;
    int j;
    for (j = 0; j < m && it < maxit; j++, it++) {
      apply_op(b, use_fmm, Vk + (size_t)j * N, w);
      for (int q = 0; q <= j; q++) {
        fo_cplx h = 0.0;
        const fo_cplx *vq = Vk + (size_t)q * N;
        for (int i = 0; i < N; i++) h += conj(vq[i]) * w[i];
        for (int i = 0; i < N; i++) w[i] -= h * vq[i];
        Hh[q + (size_t)(m + 1) * j] = h;
      }
      double hn = vnorm(N, w);
      Hh[j + 1 + (size_t)(m + 1) * j] = hn;
      if (hn > 0)
        for (int i = 0; i < N; i++) Vk[(size_t)(j + 1) * N + i] = w[i] / hn;
      for (int q = 0; q < j; q++) {
        fo_cplx h0 = Hh[q + (size_t)(m + 1) * j], h1 = Hh[q + 1 + (size_t)(m + 1) * j];
        Hh[q + (size_t)(m + 1) * j] = conj(cs[q]) * h0 + conj(sn[q]) * h1;
        Hh[q + 1 + (size_t)(m + 1) * j] = -sn[q] * h0 + cs[q] * h1;
      }
      fo_cplx a = Hh[j + (size_t)(m + 1) * j], bb = Hh[j + 1 + (size_t)(m + 1) * j];
      double den = sqrt(creal(a * conj(a)) + creal(bb * conj(bb)));
      cs[j] = a / den;
      sn[j] = bb / den;
      Hh[j + (size_t)(m + 1) * j] = den;
      Hh[j + 1 + (size_t)(m + 1) * j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = conj(cs[j]) * g[j];
      if (cabs(g[j + 1]) / bn <= tol) {
        j++;
        it++;
        break;
      }
    }
    /* back substitution */
    fo_cplx *yv = (fo_cplx *)malloc(sizeof(fo_cplx) * (j > 0 ? j : 1));
    for (int q = j - 1; q >= 0; q--) {
      fo_cplx s = g[q];
      for (int c = q + 1; c < j; c++) s -= Hh[q + (size_t)(m + 1) * c] * yv[c];
      yv[q] = s / Hh[q + (size_t)(m + 1) * q];
    }
    for (int q = 0; q < j; q++)
      for (int i = 0; i < N; i++) x[i] += yv[q] * Vk[(size_t)q * N + i];
    free(yv);
  }
  if (!conv) {
    apply_op(b, use_fmm, x, w);
    for (int i = 0; i < N; i++) r[i] = rhs[i] - w[i];
    rr = vnorm(N, r) / bn;
    conv = rr <= tol;
  }
  for (int i = 0; i < N; i++) {
    xre[i] = creal(x[i]);
    xim[i] = cimag(x[i]);
  }
  *iters = it;
  *relres = rr;
  free(rhs);
  free(x);
  free(r);
  free(w);
  free(Vk);
  free(Hh);
  free(cs);
  free(sn);
  free(g);
  return conv ? 0 : 1;
}

/* --------------------------------------------------------------- Mie */

/* p(a, theta) = sum_n (2n+1) i^n P_n(cos theta) i / ((ka)^2 h_n'(ka)) */
void fo_mie_rigid(double ka, int npts, const double *ct, int nterms,
                  double *re, double *im) {
  double *jn = (double *)malloc(sizeof(double) * (nterms + 2));
  double *yn = (double *)malloc(sizeof(double) * (nterms + 2));
  fo_sph_jn(nterms + 1, ka, jn);
  fo_sph_yn(nterms + 1, ka, yn);
  fo_cplx *cn = (fo_cplx *)malloc(sizeof(fo_cplx) * (nterms + 1));
  for (int n = 0; n <= nterms; n++) {
    fo_cplx h = jn[n] + I * yn[n];
    fo_cplx hp;
    if (n == 0)
      hp = -(jn[1] + I * yn[1]);
    else
      hp = (jn[n - 1] + I * yn[n - 1]) - (n + 1.0) / ka * h;
    cn[n] = (2.0 * n + 1.0) * cpow(I, n) * I / (ka * ka * hp);
  }
  for (int q = 0; q < npts; q++) {
    double x = ct[q], P0 = 1.0, P1 = x;
    fo_cplx s = cn[0] * P0;
    if (nterms >= 1) s += cn[1] * P1;
    for (int n = 2; n <= nterms; n++) {
      double P2 = ((2.0 * n - 1.0) * x * P1 - (n - 1.0) * P0) / n;
      s += cn[n] * P2;
      P0 = P1;
      P1 = P2;
    }
    re[q] = creal(s);
    im[q] = cimag(s);
  }
  free(jn);
  free(yn);
  free(cn);
}
