/*
 * fmb_hip.h -- the thin C-ABI layer between the C host code and the HIP
 * kernels.  Every function returns FMB_OK or a negative fmbem.h code.
 * 'stream' is a hipStream_t passed as void*.
 */
#ifndef FMB_HIP_H
#define FMB_HIP_H

#include <stddef.h>

#include "../fmb_internal.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime */
int fmbh_init(void);
int fmbh_device_tag(char *buf, int n);
int fmbh_malloc(void **p, size_t bytes);
int fmbh_free(void *p);
int fmbh_h2d(void *dst, const void *src, size_t bytes, void *stream);
int fmbh_d2h(void *dst, const void *src, size_t bytes, void *stream);
int fmbh_d2d(void *dst, const void *src, size_t bytes, void *stream);
int fmbh_memset(void *dst, int v, size_t bytes, void *stream);
int fmbh_sync(void *stream);
int fmbh_event_create(void **ev);
int fmbh_event_record(void *ev, void *stream);
int fmbh_event_elapsed(void *ev0, void *ev1, float *ms);
int fmbh_event_destroy(void *ev);
int fmbh_event_create_sync(void **ev);
int fmbh_stream_create(void **s);
int fmbh_stream_destroy(void *s);
int fmbh_stream_wait(void *stream, void *ev);
int fmbh_capture_begin(void *stream);
int fmbh_capture_end(void *stream, void **exec);
int fmbh_graph_launch(void *exec, void *stream);
int fmbh_graph_destroy(void *exec);

/* ---- tree (tree.hip) */
int fmbh_centroid_bbox(const double *d_V, const int *d_F, int n,
                       double *d_ctr, double *h_lo3, double *h_hi3,
                       void *stream);
int fmbh_morton_sort(const double *d_ctr, int n, const double *pmin,
                     double D, int *d_key10, int *d_perm, void *stream);
int fmbh_deter_lmax(const int *d_key10, int n, int s, int *lmax,
                    void *stream);
int fmbh_leaf_level(fmb_tree *t, const int *d_key10, void *stream);
int fmbh_parent_level(fmb_tree *t, int l, void *stream);
int fmbh_m2l_pairs(fmb_tree *t, int l, void *stream);
int fmbh_near_lists(fmb_tree *t, void *stream);
int fmbh_sorted_geometry(fmb_tree *t, const double *d_V, const int *d_F,
                         void *stream);

/* ---- operators (ops.hip) */
int fmbh_gen_H(const double *h_beta, int nbeta, int P, double *d_H,
               void *stream);
size_t fmbh_coax_scratch_bytes(void);
int fmbh_gen_coax_static(int kind, double k, const double *d_t, int n, int P, fmb_c64 *d_out,
                         void *stream);
int fmbh_gen_coax(int kind, double k, const double *h_t, int n, int P,
                  fmb_c64 *d_out, void *stream);
int fmbh_coax_scale(int prec, const fmb_c64 *d_in, int P, const double *d_rs,
                    const double *d_cs, int invert_rs, int invert_cs,
                    void *d_out, void *stream);
int fmbh_convert_real(int prec, const double *d_in, size_t n, void *d_out,
                      void *stream);
int fmbh_rot_phase(const double *d_H, int n, int P, const double *h_angles,
                   fmb_c64 *d_out, void *stream);

/* ---- bases and near field (basis.hip, near.hip) */
int fmbh_bases(fmb_plan *pl, int op, const fmb_c64 *d_robin_sorted,
               void *stream);
int fmbh_near_assemble(fmb_plan *pl, int op, const fmb_c64 *d_robin_sorted,
                       void *stream);
int fmbh_near_layout(fmb_plan *pl, void *stream);

/* ---- FMM pass (fmm.hip) */
int fmbh_gather(const fmb_plan *pl, int op, const fmb_c64 *x, int nrhs,
                void *stream);
int fmbh_scatter(const fmb_plan *pl, int op, fmb_c64 *y, int nrhs,
                 void *stream);
int fmbh_p2m(const fmb_plan *pl, int op, int nrhs, void *stream);
int fmbh_m2m(const fmb_plan *pl, int op, int l, int nrhs, void *stream);
int fmbh_m2l(const fmb_plan *pl, int op, int l, int nrhs, void *stream);
/* matrix-core RCR translations (rcr_kernels.h, rcr_mfma.hip) */
#define FMBH_RCR_LDS 1 /* batch in LDS, blocks owned by one wave          */
#define FMBH_RCR_GWS 2 /* batch in a global workspace, row-group units    */
#define FMBH_RCR_LDSP 3 /* as LDS, a persistent grid looping over batches */
#define FMBH_RCR_LDSO 4 /* as LDSP, with the bucket's operators in LDS      */
typedef struct fmbh_rcr_job {
  int mode, nt, mt, nw; /* nw: waves per workgroup, 4 or 8 */
  int bf;               /* split-BF16 rotations (FP32 pass, LDS modes)   */
  const void *Hb;       /* their operator blocks, per beta: hi then lo    */
  size_t hbstride, hbhalf;
  const void *chunks;
  int nchunks, chunk_pairs;
  const void *info;
  const void *phase; /* per bucket: e^{i m phi}, m < pstride */
  int pstride;
  const int *tgt, *src;
  const void *H;
  size_t hstride;
  const void *C;
  size_t cstride;
  int P, pin, pout;
  const int *sched;
  int sched_len;
  int hlen;          /* values of the padded rotation blocks n < P        */
  int xcd;           /* 1: LDS mode on a 1-D grid, each XCD taking a
                        contiguous range of the (chunk, batch) items      */
  const void *in;
  void *out;
  size_t in_bytes, out_bytes; /* extents of in and out this pass reaches */
  void *gws;
  int gws_grid;
} fmbh_rcr_job;
/* workspace mode and tile shape for max(pin, pout) = pm (the LDS budget
 * leaves room for a schedule of up to 8 pm + 16 ints) */
int fmbh_rcr_shape(int pm, int prec, int *mode, int *nt, int *mt);
/* global workspace bytes of one workgroup in GWS mode */
size_t fmbh_rcr_lds_bytes(int pm, int prec, int nt, int sched_len);
int fmbh_rcr_mt_lds(int pm, int bf);
size_t fmbh_rcr_gws_bytes(int pm, int prec, int nt);
/* LDSO-mode bytes of a workgroup (0: no LDS-mode unit shape for pm) */
size_t fmbh_rcr_ldso_bytes(int pm, int prec, int nt, int sched_len, int hlen, size_t clen);
int fmbh_rcr(const fmbh_rcr_job *j, int prec, int nrhs, void *stream);
/* zero-padded operator copies: degree n's even/odd rotation blocks (from
 * the FP64 blocks, rounded to the opset's type) at hoff[n], the even block
 * ((n+1) x (n+1), index i for m = i + 1, i < n, and m = 0) then the odd one
 * (n x n, index i for |m| = i + 1), each to 16-row x 4-column multiples,
 * leading dimension = padded rows; coaxial block |m| (rows pout-|m|,
 * columns pin-|m|) at coff[|m|] */
int fmbh_pad_rot(int prec, const double *d_H64, int nbeta, int Pmax,
                 const int *d_hoff, size_t pstride, void *d_Hp, void *stream);
/* split-BF16 copies of the standard rotation blocks H_n of beta b (from
 * the FP64 blocks): at hboff[n] (rows padded to 16, K to 32, row-major), hi
 * part at b * 2 half, lo part half values further */
int fmbh_split_rot(const double *d_H64, int nbeta, int Pmax, const int *d_hboff, size_t half,
                   void *d_Hb, void *stream);
int fmbh_pad_coax(int prec, const void *d_C, int ncoax, int P, int pin,
                  int pout, const int *d_coff, size_t pstride, void *d_Cp,
                  void *stream);
int fmbh_l2l(const fmb_plan *pl, int op, int l, int nrhs, void *stream);
int fmbh_l2p_near(const fmb_plan *pl, int op, int nrhs, int parts,
                  void *stream);
int fmbh_mmajor_table(int p, int **d_tab, void *stream);

/* ---- FP64 vector kernels for GMRES (blas.hip) */
/* pressure at field points (field.hip): d_pts npts*3, d_dirs nrhs*3 or NULL */
int fmbh_field(const fmb_plan *pl, const fmb_c64 *x, int nrhs, const double *d_pts,
               int npts, const double *d_dirs, fmb_c64 *out, void *stream);
/* the same through the tree, from the multipole expansions that opset op
 * (FP64) holds for the nrhs vectors of x (field.hip); d_pts in Morton
 * order, d_qperm[i] the caller's index of point i */
int fmbh_field_tree(const fmb_plan *pl, int op, const fmb_c64 *x, int nrhs,
                    const double *d_pts, const int *d_qperm, int npts,
                    const double *d_dirs, fmb_c64 *out, void *stream);
/* transposed pass pieces: (-1)^n L2P^T into L[lmax]; near^T + P2M^T into
 * ys with the given near blocks / P2M basis (NULL: the opset's own) */
int fmbh_l2p_t(const fmb_plan *pl, int op, int nrhs, void *stream);
int fmbh_near_t(const fmb_plan *pl, int op, const void *near, const void *basis,
                int nrhs, int parts, void *stream);
/* sigma-derivative tables for the adjoint: P2M basis of dA/dsigma_j's
 * column factor ([N][pL^2]) and its near blocks (near layout) */
int fmbh_bases_dsigma(fmb_plan *pl, int op, void *out, void *stream);
int fmbh_near_dsigma(fmb_plan *pl, int op, void *out, void *stream);
/* transposed field evaluation: kt, st [nrhs][N] caller order from weights
 * u [nrhs][npts] (field.hip) */
int fmbh_field_t(const fmb_plan *pl, const fmb_c64 *u, int nrhs, const double *d_pts,
                 int npts, fmb_c64 *kt, fmb_c64 *st, void *stream);
int fmbh_adjoint_combine(const fmb_c64 *x, const fmb_c64 *c, const fmb_c64 *s,
                         int nrhs, int N, fmb_c64 *g, void *stream);
/* vertex gradients of the field objective (shape.hip) */
int fmbh_shape_grad(const fmb_plan *pl, const fmb_c64 *x, const fmb_c64 *lam, int nrhs,
                    const void *Lx, const void *Mt, const fmb_c64 *u, const double *d_pts,
                    int npts, const double *d_dirs, double *dV, double *dvn, void *stream);
int fmbh_sweep_guess(const fmb_plan *pl, double dk, const fmb_c64 *d_x,
                     const double *h_dirs, int nrhs, fmb_c64 *d_x0, void *stream);
int fmbh_rhs_plane(const fmb_plan *pl, const double *h_dirs, int nrhs,
                   fmb_c64 *d_b, void *stream);
/* h[j*nrhs + r] = <V_j^r, w^r> for j < nv ; V: [nv][nrhs][N] */
int fmbh_zdots(const fmb_c64 *d_V, int nv, const fmb_c64 *d_w, int nrhs,
               int N, fmb_c64 *d_h, void *stream);
/* w^r -= sum_j h[j*nrhs+r] V_j^r */
int fmbh_zgemv_sub(const fmb_c64 *d_V, int nv, const fmb_c64 *d_h,
                   fmb_c64 *d_w, int nrhs, int N, void *stream);
/* nrm[r] = ||w^r|| */
int fmbh_znorms(const fmb_c64 *d_w, int nrhs, int N, double *d_nrm,
                void *stream);
/* y^r = a[r] x^r + b[r] y^r  (a, b device complex per rhs; NULL b = 0) */
int fmbh_zaxpby(const fmb_c64 *d_a, const fmb_c64 *d_x, const fmb_c64 *d_b,
                fmb_c64 *d_y, int nrhs, int N, void *stream);
/* x^r += sum_j y[j*nrhs+r] V_j^r  (y on device) */
int fmbh_zgemv_add(const fmb_c64 *d_V, int nv, const fmb_c64 *d_y,
                   fmb_c64 *d_x, int nrhs, int N, void *stream);

#ifdef __cplusplus
}
#endif
#endif
