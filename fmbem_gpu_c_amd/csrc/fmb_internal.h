/*
 * fmb_internal.h -- data structures shared by the C host code (csrc/host)
 * and the HIP layer (csrc/hip).  Plain C: device buffers are void* or
 * typed pointers into device memory.
 */
#ifndef FMB_INTERNAL_H
#define FMB_INTERNAL_H

#include <stddef.h>

#include "../../include/fmbem.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FMB_MAXLEV 12 /* levels 0..11; the reference caps at 10 (octree.c:637) */
#define FMB_NOFF 343  /* offset codes (dx+3)*49+(dy+3)*7+(dz+3)               */
#define FMB_NDIST 15  /* distinct |offset|^2 among M2L offsets                */

struct fmb_mesh {
  int nv, nf;
  double *V; /* nv*3 */
  int *F;    /* nf*3 */
};

struct fmb_tree {
  int nelem;
  int lmin, lmax;
  double D, pmin[3];
  int nbox[FMB_MAXLEV];
  /* device arrays */
  int *d_keys[FMB_MAXLEV];   /* sorted Morton keys                       */
  int *d_parent[FMB_MAXLEV]; /* l > 2: parent index in level l-1          */
  int *d_cstart[FMB_MAXLEV]; /* l < lmax: children range in level l+1, nbox+1 */
  int *d_perm;               /* sorted position -> caller element id       */
  int *d_iperm;              /* caller element id -> sorted position       */
  int *d_box_start;          /* leaf: nbox+1 element ranges (sorted order) */
  int *d_elem_box;           /* sorted position -> leaf box                */
  double *d_tri;             /* sorted order: 9 doubles per element        */
  double *d_geo;             /* sorted order: centroid(3) normal(3) area radius */
  int *d_fv;                 /* sorted order: the element's 3 vertex ids    */
  int nvert;
  double *d_vnorm;           /* [nvert][3] unit area-weighted vertex normals */
  int npairs[FMB_MAXLEV];
  int *d_pair_tgt[FMB_MAXLEV], *d_pair_src[FMB_MAXLEV];
  int *h_off_start[FMB_MAXLEV]; /* host: FMB_NOFF+1 bucket ranges           */
  int *d_near_start, *d_near_box; /* leaf: near boxes incl. self (CSR)     */
  int near_nnzbox;
  int max_leaf_elems;
  void *stream;
};

/* one precision's operators, bases, near field and workspace */
typedef struct fmb_opset {
  int prec;      /* FMB_FP64 or FMB_MIXED(=FP32 arithmetic)                 */
  size_t csz;    /* bytes of one complex value                              */
  void *d_H;     /* [nbeta][rot_size(Pmax)] real                            */
  void *d_Csr[FMB_MAXLEV];  /* [15][coax_size(p_l)] complex, scaled           */
  void *d_Cm2m[FMB_MAXLEV]; /* l>2: (R|R) child l -> parent l-1, scaled       */
  void *d_Cl2l[FMB_MAXLEV]; /* l>2: (R|R) parent l-1 -> child l, scaled       */
  void *d_P2M;  /* [N][pL^2] complex, sorted order, scaled                   */
  void *d_L2P;  /* [N][pL^2] complex                                        */
  void *d_near; /* near blocks, complex                                     */
  void *d_M[FMB_MAXLEV], *d_L[FMB_MAXLEV]; /* [nbox][nrhs][p^2] complex      */
  void *d_xs, *d_ys; /* sorted-order vectors [nrhs][N] complex                */
  void *d_gws;       /* global workspace of the GWS-mode translations: one
                        slot per stream the plan launches them on         */
  size_t gws_bytes, gws_slot;
  void *d_Hpad;      /* zero-padded rotation blocks [nbeta][hpad_stride]    */
  void *d_phase[3];  /* e^{i m phi}, m < Pmax: M2L offsets, M2M, L2L octants */
  void *d_phase_c[3]; /* e^{-i m phi}: the same buckets, transposed pass    */
  size_t hpad_stride;
  void *d_Hb;        /* FP32 pass with split_bf16: rotation blocks as bf16
                        hi and lo parts [nbeta][2][hb_half]                 */
  size_t hb_half;
  int gws_grid;      /* persistent workgroups sharing d_gws                  */
} fmb_opset;

/* one translation kind at one level on the matrix cores */
typedef struct fmb_rcr_tab {
  int *d_par, *d_chi;   /* M2M/L2L: (parent, child) pairs by child octant */
  void *d_chunks;       /* int4 {bucket, first pair, end pair, 0}          */
  int nchunks;
  int pin, pout, P;     /* source order, target order, coaxial order       */
  int mode[2], nt[2], mt[2]; /* per opset: 0 = VALU kernel, else FMBH_RCR_* */
  int nw[2];            /* per opset: waves per workgroup (4 or 8)         */
  int bf[2];            /* per opset: split-BF16 rotations                 */
  int *d_sched[2];      /* per opset: unit schedule                        */
  int sched_len[2];     /* ints in d_sched                                 */
  void *d_Cpad[2];      /* per opset: zero-padded coaxial blocks           */
  float tuned_ms[2];    /* per opset: autotuned time of one launch (0: not tuned) */
  size_t cpad_stride;   /* complex values per coaxial matrix in d_Cpad     */
  int hlen;             /* values of the padded rotation blocks n < P      */
  int tiled;            /* M2L: chunks in tile-major order (else offset-major) */
} fmb_rcr_tab;

struct fmb_plan {
  const fmb_tree *tree;
  fmb_plan_params prm;
  double k, alpha_re, alpha_im;
  int p[FMB_MAXLEV];
  int Pmax;  /* order of the rotation tables: >= every level's order        */
  double *h_scale[FMB_MAXLEV]; /* w_n per level                              */
  double *d_scale[FMB_MAXLEV];
  int nbeta;
  double beta_cos[FMB_NOFF];
  int off_beta[FMB_NOFF];
  double off_phi[FMB_NOFF];
  int off_dist[FMB_NOFF];
  /* octant o of a child: child centre - parent centre = (bits - 1/2) side */
  int m2m_beta[8], l2l_beta[8]; /* t = parent - child, t = child - parent  */
  double m2m_phi[8], l2l_phi[8];
  int *d_m2m_beta, *d_l2l_beta;
  double *d_m2m_phi, *d_l2l_phi;
  void *d_off_info;     /* FMB_NOFF x {int beta, int dist, double phi}      */
  void *d_m2l_chunks[FMB_MAXLEV]; /* int4 {offset, first, last, 0}          */
  int n_m2l_chunks[FMB_MAXLEV];
  /* matrix-core translations, [kind][level]: kind 0 M2L at l, 1 M2M
   * l -> l-1, 2 L2L l-1 -> l (csrc/host/rcr.c) */
  fmb_rcr_tab rcr[3][FMB_MAXLEV];
  void *d_info_oct[2]; /* M2M, L2L: per child octant {beta, 0, phi}        */
  double *d_quad;       /* far rule: nq x {b0,b1,b2,w}                    */
  double *d_quadn;      /* close rule                                      */
  double *d_edge;       /* nedge x {x, w} Gauss-Legendre on [-1,1]         */
  int nq, nqn;
  size_t *d_near_off;   /* per leaf box: offset of its block (elements)    */
  size_t near_total;
  int *d_near_ncol;     /* per leaf box: columns of its block              */
  int near_max_ncol;
  int *d_tab_mmajor[FMB_MAXLEV]; /* per level: m-major (n',m) enumeration   */
  fmb_opset ops[2];     /* [0] FP64, [1] FP32 (FMB_MIXED only)            */
  int nops;
  int work_prec;        /* precision of the Arnoldi matvecs in fmb_gmres   */
  int nrhs_max;
  double *d_work64;     /* [nrhs][N] FP64 complex scratch                  */
  fmb_c64 *d_robin;     /* Robin coefficients in Morton order, or NULL     */
  int timing;
  double t_phase[7];
  double t_m2l[FMB_MAXLEV];
  size_t bytes;         /* device bytes of the rotation + frequency tiers  */
  size_t bytes_rot, bytes_freq;
  int ready;            /* frequency tier built (cleared by a failed retune) */
  /* a second stream for the coarse M2L levels and the L2L chain, and the
   * events ordering it against the fine M2L levels (NULL: one stream) */
  void *aux, *ev_up, *ev_aux, *ev_m2l[FMB_MAXLEV];
  int capturing;        /* set while a batch captures this plan's pass      */
  void *stream;
};

/* offset code helpers */
static inline int fmb_off_code(int dx, int dy, int dz) {
  return (dx + 3) * 49 + (dy + 3) * 7 + (dz + 3);
}

int fmb_set_error(int code, const char *fmt, ...);

#ifdef __cplusplus
}
#endif
#endif
