/*
 * fmb_oracle.h -- FP64 CPU restatement of the FMM-BEM algorithm of the
 * reference FMBEM_GPU_C (octree.c, translation.cu, integral.cu, structs.cu).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * in fmbem_gpu_c_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * Conventions (pinned by the tests under tests/):
 *   coefficient index  (n,m) -> n*n+n+m                 (translation.h:60-62)
 *   Y_n^m = (-1)^m sqrt((2n+1)/4pi (n-|m|)!/(n+|m|)!) P_n^|m|(cos t) e^{i m phi}
 *           with P_n^|m| carrying the Condon-Shortley phase, i.e. the product
 *           has no (-1)^m left                           (translation.cu:153-169)
 *   R_n^m = j_n(kr) Y_n^m,  S_n^m = h_n(kr) Y_n^m        (translation.cu:171-187)
 *   G = e^{ikr}/(4 pi r)                                 (integral.cu:72-78)
 *   sparse rotation / coaxial layouts                    (translation.cu:12-56)
 *
 * Parity status per function is listed in the header comment of each
 * section of fmb_oracle.c.
 */
#ifndef FMB_ORACLE_H
#define FMB_ORACLE_H

#include <complex.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex fo_cplx;

/* ---------------- special functions ---------------- */
double fo_acoef(int n, int m);
double fo_bcoef(int n, int m);
void fo_sph_jn(int nmax, double x, double *j);
void fo_sph_yn(int nmax, double x, double *y);
void fo_legendre_table(int nmax, double ct, double st, double *P);
void fo_ylm(int nmax, double theta, double phi, double *re, double *im);
void fo_rbasis(int nmax, double k, const double *v, double *re, double *im);
void fo_sbasis(int nmax, double k, const double *v, double *re, double *im);
void fo_rbasis_dn(int nmax, double k, const double *v, const double *nrm,
                  double *re, double *im);

/* ---------------- quadrature and panel integrals ---------------- */
#define FO_RULE_COLLAPSED 0 /* INTORDER^2 collapsed rule, integral.cu:80-139 */
#define FO_RULE_GAUSS7 1    /* 7-point degree-5 rule (Strang-Fix)          */
int fo_gauss_legendre(int n, double *x, double *w);
int fo_tri_rule(int rule, int order, double *bary, double *w);
void fo_kernels(double k, const double *x, const double *nx, const double *y,
                const double *ny, double *re4, double *im4);
void fo_panel_regular(double k, const double *tri, const double *x,
                      const double *nx, int rule, int order, double *re4,
                      double *im4);
void fo_panel_self(double k, const double *tri, int nedge, double *re4,
                   double *im4);
void fo_panel_hyper_ref(double k, const double *tri, int order, double *re,
                        double *im);

/* ---------------- octree (octree.c restatement) ---------------- */
int fo_trunc_num(double k, double eps, double sigma, double a);
int fo_morton_key(const double *s, int l);
void fo_box_center(int key, int l, double *c);
int fo_neighbors(int key, int l, int *nbr);
int fo_I4(int key, int l, int *set);
int fo_deter_lmax(const double *pts, int n, int s);
int fo_bounding_cube(const double *pts, int n, double eps, double *pmin,
                     double *d);

typedef struct fo_tree {
  int nelem;
  int lmin, lmax;
  double D, pmin[3];
  int nbox[16];    /* boxes per level (index = level)          */
  int *keys[16];   /* sorted Morton keys per level            */
  int *perm;       /* elements sorted by leaf key: perm[i] = original id */
  int *elem_box;   /* leaf-box index of each original element  */
  int *box_start;  /* leaf level: element range in perm order  */
  int *box_end;
} fo_tree;

int fo_tree_build(const double *V, int nv, const int *F, int nf, int s,
                  int depth, fo_tree **out);
void fo_tree_free(fo_tree *t);
int fo_tree_lmax(const fo_tree *t);
int fo_tree_nbox(const fo_tree *t, int l);
void fo_tree_keys(const fo_tree *t, int l, int *keys);
void fo_tree_geom(const fo_tree *t, double *D, double *pmin);
void fo_tree_perm(const fo_tree *t, int *perm);
int fo_tree_num_sr(const fo_tree *t, int l, int *num);

/* ---------------- translation operators (translation.cu) ---------------- */
int fo_rot_size(int p);
int fo_coax_size(int p);
int fo_coax_block(int p, int m);
void fo_rot_H(int p, double beta, double *H);
void fo_rot_H_ref(int p, double beta, double *H);
void fo_rot_sparse(int p, double alpha, double beta, double gamma,
                   double *re, double *im);
#define FO_COAX_RR 0
#define FO_COAX_SR 1
void fo_coax_sparse(int kind, int p, double k, double t, double *re,
                    double *im);
void fo_rot_apply(int p, const double *mre, const double *mim,
                  const double *xre, const double *xim, double *yre,
                  double *yim);
void fo_coax_apply(int P, int pin, int pout, const double *mre,
                   const double *mim, const double *xre, const double *xim,
                   double *yre, double *yim);
void fo_translate(int kind, int pin, int pout, double k, const double *t,
                  const double *xre, const double *xim, double *yre,
                  double *yim);

/* ---------------- BEM (Burton-Miller, constant elements) ---------------- */
typedef struct fo_bem fo_bem;
int fo_bem_create(const double *V, int nv, const int *F, int nf, double k,
                  const double *alpha, /* coupling (re,im)              */
                  const double *sig_re, const double *sig_im, /* dp/dn = sig p */
                  int rule, int order, int nedge, fo_bem **out);
void fo_bem_free(fo_bem *b);
void fo_bem_set_near(fo_bem *b, int rule, int order, double ratio);
int fo_bem_n(const fo_bem *b);
void fo_bem_dense_row(const fo_bem *b, int i, double *re, double *im);
void fo_bem_dense_matvec(const fo_bem *b, int nrhs, const double *xre,
                         const double *xim, double *yre, double *yim);
void fo_bem_rhs_plane(const fo_bem *b, const double *dir, double *re,
                      double *im);
void fo_bem_centroids(const fo_bem *b, double *c, double *nrm, double *area);
/* Pressure at field points (exterior representation formula): x [nrhs][N]
 * surface solution; pts npts*3; dirs nrhs*3 incident plane waves or NULL;
 * out [nrhs][npts]. */
void fo_bem_field(const fo_bem *b, int nrhs, const double *xre, const double *xim,
                  const double *pts, int npts, const double *dirs, double *ore,
                  double *oim);

int fo_bem_fmm_setup(fo_bem *b, int s, int depth, const int *plev);
int fo_bem_fmm_orders(const fo_bem *b, int *p);
void fo_bem_fmm_matvec(const fo_bem *b, int nrhs, const double *xre,
                       const double *xim, double *yre, double *yim);
/* Pressure at field points through the tree: the multipole expansions of
 * the boxes well separated from each point, the panels of the leaf boxes
 * next to it by the rule of fo_bem_field (needs fo_bem_fmm_setup). */
void fo_bem_field_fmm(const fo_bem *b, int nrhs, const double *xre,
                      const double *xim, const double *pts, int npts,
                      const double *dirs, double *ore, double *oim);
void fo_bem_near_matvec(const fo_bem *b, int nrhs, const double *xre,
                        const double *xim, double *yre, double *yim);

/* GMRES(m) with FP64 modified Gram-Schmidt; use_fmm selects the operator */
int fo_gmres(const fo_bem *b, int use_fmm, int m, double tol, int maxit,
             const double *bre, const double *bim, double *xre, double *xim,
             int *iters, double *relres);

/* rigid-sphere Mie series: total surface pressure, plane wave e^{ik d.x} */
void fo_mie_rigid(double ka, int npts, const double *cos_theta, int nterms,
                  double *re, double *im);

int fo_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
