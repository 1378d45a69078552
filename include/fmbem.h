/*
 * fmbem.h -- C ABI of the MI355X (gfx950) fast multipole boundary element
 * solver for 3-D Helmholtz acoustics.
 *
 * Every entry point has C linkage, takes plain pointers and sizes, and
 * returns FMB_OK or a negative error code (text via fmb_last_error()).
 * Streams are hipStream_t passed as void* (NULL = the default stream).
 * Vectors handed to fmb_matvec / fmb_gmres / fmb_rhs_plane_wave live in
 * device memory, as nrhs consecutive vectors of N complex FP64 values in the
 * caller's element order.  Handles are thread-compatible: one handle per
 * host thread at a time.
 *
 * Each declaration names the reference interface it replaces
 * (ZiqiFrancisFan/FMBEM_GPU_C, file:line).
 */
#ifndef FMBEM_H
#define FMBEM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ basics */
#define FMB_OK 0
#define FMB_ERR_ARG (-1)     /* invalid argument                          */
#define FMB_ERR_IO (-2)      /* file missing or malformed                 */
#define FMB_ERR_HIP (-3)     /* HIP runtime error                         */
#define FMB_ERR_NOMEM (-4)   /* allocation failed                         */
#define FMB_ERR_DEPTH (-5)   /* s+1 centroids share a level-10 box        */
#define FMB_ERR_NOCONV (-6)  /* GMRES did not reach the tolerance         */
#define FMB_ERR_RANGE (-7)   /* order / size beyond a supported limit     */

typedef struct { float re, im; } fmb_c32;
typedef struct { double re, im; } fmb_c64;

/* Replaces the HOST_CALL/CUDA_CALL "Error at file:line" prints
 * (translation.h:64-126): message of the last failing call on this thread. */
const char *fmb_last_error(void);
int fmb_version(void);

/* ------------------------------------------------------------ mesh */
typedef struct fmb_mesh fmb_mesh;

/* findNum (octree.c:563-584; octree.h:30): counts 'v' and 'f' lines. */
int fmb_obj_count(const char *path, int *nv, int *nf);
/* readOBJ (octree.c:586-619; octree.h:32): 'v x y z', 'f a b c' (1-based). */
int fmb_mesh_load_obj(const char *path, fmb_mesh **out);
/* Mesh from caller arrays: V[nv*3] doubles, F[nf*3] 0-based ids. */
int fmb_mesh_from_arrays(const double *V, int nv, const int *F, int nf,
                         fmb_mesh **out);
/* Icosphere with a vertex at each pole; level 3, radius 0.1 gives the vertex
 * set of the reference's sphere_100mm.obj.  Faces are wound so that the
 * reference normal (n0-n2)x(n1-n2) (integral.cu:420) points outward. */
int fmb_mesh_icosphere(int level, double radius, fmb_mesh **out);
/* Ellipsoid (semi-axes a,b,c) from a geodesic sphere of frequency nu
 * (20 nu^2 faces), radially perturbed by a real spherical-harmonic field of
 * degrees lmin..lmax with N(0,1) coefficients (seeded) scaled to rms_frac
 * of the radius. */
int fmb_mesh_ellipsoid(int nu, double a, double b, double c, int lmin,
                       int lmax, double rms_frac, unsigned long long seed,
                       fmb_mesh **out);
/* Re-winds faces so the reference normal points away from the enclosed
 * volume (the shipped sphere meshes point inward). Returns faces flipped. */
int fmb_mesh_orient_outward(fmb_mesh *m, int *nflipped);
int fmb_mesh_sizes(const fmb_mesh *m, int *nv, int *nf);
int fmb_mesh_export(const fmb_mesh *m, double *V, int *F);
void fmb_mesh_destroy(fmb_mesh *m);

/* ------------------------------------------------------------ octree */
typedef struct fmb_tree fmb_tree;

/* srcBoxes + FMMLevelSet + FMMLevelSetNumSR + the index arrays of genOctree
 * (octree.c:749-768, 890-919; translation.cu:3007-3076, 2797-2969), built on
 * the device as Morton-sorted arrays.  s: max elements per leaf box
 * (deterLmax, octree.c:635-679) used when depth == 0; depth > 0 fixes the
 * leaf level.  lmin is 2 (translation.cu:2976). */
int fmb_tree_build(const fmb_mesh *m, int s, int depth, void *stream,
                   fmb_tree **out);
/* lmax, D (cube side), pmin (cube corner, 3 doubles) (struct octree,
 * structs.h:93-150). */
int fmb_tree_info(const fmb_tree *t, int *lmin, int *lmax, double *D,
                  double *pmin);
/* FMMLevelSet (octree.c:890-906): sorted Morton keys of the non-empty boxes
 * at level l; keys may be NULL to query the count. */
int fmb_tree_level(const fmb_tree *t, int l, int *nbox, int *keys);
/* FMMLevelSetNumSR (octree.c:908-919): |I4(b) & level set| per box. */
int fmb_tree_num_sr(const fmb_tree *t, int l, int *num);
/* btmLvlElemIdx (translation.cu:3062-3076): leaf-box index per element, and
 * the Morton order of the elements (perm[i] = original id of sorted i). */
int fmb_tree_elem_box(const fmb_tree *t, int *box, int *perm);
/* M2L pairs of level l (srTransOriginArr / srTransDestArr,
 * translation.cu:2847-2927) as box indices, offset-major order; offsets are
 * codes (dx+3)*49+(dy+3)*7+(dz+3) of target-minus-source grid steps. */
int fmb_tree_pairs(const fmb_tree *t, int l, int *npairs, int *tgt, int *src,
                   int *off);
/* Unit vertex normals (nvert*3 host doubles): normalised sums of the
 * adjacent faces' (n0-n2)x(n1-n2), i.e. area-weighted face normals.  The
 * reference has none (used by fmb_adjoint_grad's dJ_dvn). */
int fmb_tree_vertex_normals(const fmb_tree *t, double *normals);
void fmb_tree_destroy(fmb_tree *t);

/* ------------------------------------------------------------ operators */
/* truncNum (octree.c:777-796; octree.h:112). */
int fmb_trunc_num(double k, double eps, double sigma, double a);
int fmb_rot_size(int p);  /* sparseRotSize, translation.cu:12-15       */
int fmb_coax_size(int p); /* sparseCoaxTransSize, translation.cu:17-20 */
/* genSparseRotMats (translation.cu:1297-1394; translation.h:224): angles is
 * n triples (alpha, beta, gamma); out (host) receives n*fmb_rot_size(p)
 * complex values in the reference n-block layout.  Generated on the device
 * in FP64 with O(p^3) work and memory. */
int fmb_gen_rot_mats(const double *angles, int n, int p, fmb_c64 *out);
/* genSRSparseCoaxTransMat / genRRSparseCoaxTransMat (translation.cu:982-1062,
 * 822-903; translation.h:174-187): kind 0 = (R|R), 1 = (S|R); n distances t,
 * out (host) receives n*fmb_coax_size(p) values in the |m|-block layout. */
#define FMB_COAX_RR 0
#define FMB_COAX_SR 1
int fmb_gen_coax_mats(int kind, double k, const double *t, int n, int p,
                      fmb_c64 *out);
/* No counterpart in the reference: fmb_gen_coax_mats for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  d_t: n distances in device memory;
 * d_out: n*fmb_coax_size(p) complex values in device memory, written on
 * `stream` without a host synchronisation.  The recurrence table
 * (n * p * (2p-1) * 16 bytes) lives in a fixed device scratch array of
 * fmb_coax_scratch_bytes() bytes per device (8 MiB, held until the process
 * ends); larger requests return FMB_ERR_RANGE.  No allocation is made.
 * Calls that overlap on one device share the scratch and race. */
int fmb_gen_coax_mats_noalloc(int kind, double k, const double *d_t, int n, int p,
                              fmb_c64 *d_out, void *stream);
size_t fmb_coax_scratch_bytes(void);
/* reduceSparseRotMat (translation.cu:22-26; translation.h:132): a rotation
 * matrix of order p1 (host, reference n-block layout) truncated to order
 * p2 <= p1: its first fmb_rot_size(p2) values (in and out may alias). */
int fmb_reduce_rot_mat(const fmb_c64 *in, int p1, fmb_c64 *out, int p2);
/* reduceSparseCoaxTransMat (translation.cu:34-56; translation.h:134): a coaxial
 * matrix of order p1 (host, |m|-block layout) truncated to order p2 <= p1:
 * block m < p2 keeps its leading (p2-m) x (p2-m) entries. */
int fmb_reduce_coax_mat(const fmb_c64 *in, int p1, fmb_c64 *out, int p2);

/* ------------------------------------------------------------ plan */
#define FMB_FP64 0  /* plain path: FP64 everywhere                       */
#define FMB_MIXED 1 /* FP32 translations / near field / bases, FP64
                       recurrences, singular terms, Arnoldi, refinement  */

#define FMB_PRULE_TRUNCNUM 0 /* truncNum(k, eps, sigma, side) per level    */
#define FMB_PRULE_EXCESS 1   /* ka_box + 1.8 log10(ka_box + pi), ka_box =
                                k * box diameter (sqrt(3) * side)          */
#define FMB_PRULE_EXPLICIT 2 /* p_levels[l] as given                       */

#define FMB_RULE_COLLAPSED 0 /* INTORDER^2 collapsed rule, integral.cu:80-139 */
#define FMB_RULE_GAUSS7 1    /* 7-point degree-5 triangle rule              */

typedef struct fmb_plan fmb_plan;

/* kernel family of the M2L / M2M / L2L translations */
#define FMB_RCR_AUTO 0 /* matrix cores (batched per offset / octant)      */
#define FMB_RCR_VALU 1 /* one translation per workgroup on the vector ALU */

/* workspace mode of the matrix-core translations (fmb_plan_params.rcr_mode) */
#define FMB_RCR_MODE_RULE 0 /* the tile-shape rule (or the autotuner)      */
#define FMB_RCR_MODE_LDS 1  /* batch in LDS, one workgroup per batch       */
#define FMB_RCR_MODE_GWS 2  /* batch in a global workspace, persistent grid */
#define FMB_RCR_MODE_LDSP 3 /* batch in LDS, persistent grid               */
#define FMB_RCR_MODE_LDSO 4 /* as LDSP, with the bucket's operators in LDS  */

typedef struct fmb_plan_params {
  double k;                  /* wavenumber (rad/m)                          */
  double alpha_re, alpha_im; /* Burton-Miller coupling; 0,0 selects i/k     */
  int precision;             /* FMB_FP64 or FMB_MIXED                       */
  int p_rule;                /* FMB_PRULE_*                                 */
  double eps, sigma;         /* truncNum parameters (0.05, 1.5)             */
  int p_levels[16];          /* FMB_PRULE_EXPLICIT: order at level l        */
  int p_min;                 /* lower clamp on every level's order          */
  int quad_rule, quad_order; /* panel rule for far pairs and P2M            */
  int near_rule, near_order; /* refined rule for close pairs                */
  double near_ratio;         /* close: |x_i - c_j| < ratio * circumradius_j */
  int nedge;                 /* Gauss points per edge, singular self terms  */
  int nrhs_max;              /* right-hand sides per matvec                 */
  int rcr_path;              /* FMB_RCR_*: translation kernel family        */
  double k_max;              /* largest wavenumber of a frequency sweep: the
                                rotation tables are built for its orders
                                once (initOctree's maxWavNum,
                                translation.cu:2992); 0: k                  */
  int autotune;              /* 1: time every valid tile shape of every
                                matrix-core translation at plan creation
                                and keep the fastest (fmb_plan_tile_shape)  */
  int serial_levels;         /* 1: M2L levels one after another on the
                                caller's stream; 0: the coarse levels and
                                the L2L chain on a second stream, beside
                                the M2L of levels L-2 and L-1 (the finest
                                M2L then runs alone)                        */
  int waves;                 /* waves per workgroup of the matrix-core
                                translations in the fixed tile-shape rule:
                                4 or 8 (0: 4); autotune overrides it        */
  int split_bf16;            /* FMB_MIXED plans: 1 runs the rotations of the
                                FP32 pass's LDS-mode translations in
                                split-BF16 (hi + lo bf16 operands, three
                                products, FP32 accumulation; about 2^-16
                                relative per product)                       */
  int ldso;                  /* 1: in the fixed tile-shape rule, every
                                translation whose batch and operator blocks
                                fit in LDS together (FP32 orders up to about
                                20) keeps the operators there on a
                                persistent grid; autotune tries that shape
                                either way                                  */
  int m2l_tile;              /* 0: M2L work in offset-major order (all pairs
                                of one offset, then the next); t > 0: tile-
                                major order (the pairs of t consecutive
                                target boxes for every offset, then the next
                                tile), each XCD taking a contiguous range of
                                tiles, so that a tile's sources and targets
                                stay in its caches, on the levels whose FP64
                                source coefficients (boxes x nrhs_max x p^2
                                x 16 B) exceed m2l_tile_min_mb; offset-major
                                on the others (default 256)                 */
  int m2l_tile_min_mb;       /* the size above which a level's M2L runs tile-
                                major, in MiB (default 256, the GPU's last-
                                level cache; 0: every level)               */
  int rcr_mode;              /* FMB_RCR_MODE_*: the fixed tile-shape rule's
                                workspace mode for every translation where
                                its batch (and for LDSO its operators) fit
                                in LDS, else the rule's own; GWS always
                                fits.  For tests and tuning studies (0: the
                                rule; the autotuner times every mode) */
} fmb_plan_params;

void fmb_plan_params_default(fmb_plan_params *prm);

/* genOctree's operator part + initOctree defaults (translation.cu:2971-3348):
 * per-level orders, rotation matrices per unique polar angle, (S|R) and
 * (R|R) coaxial matrices, P2M/L2P bases, near-field blocks (regular rule,
 * refined close rule, analytic singular self terms).  robin (device or NULL)
 * gives sigma per element (caller order) for dp/dn = sigma p; NULL: rigid. */
int fmb_plan_create(const fmb_tree *t, const fmb_plan_params *prm,
                    const fmb_c64 *robin, void *stream, fmb_plan **out);
int fmb_plan_orders(const fmb_plan *pl, int *p16);

/* genOctree called again on a filled octree (translation.cu:3007-3186,
 * 3294-3302): re-targets the plan at wavenumber k for a frequency sweep.
 * Kept: the tree, translation directions, quadrature rules, near-field layout
 * and the rotation tables (rebuilt only when an order exceeds the largest
 * built so far, or fmb_plan_params.k_max's).  Rebuilt: orders, scale
 * factors, coaxial matrices, P2M/L2P bases, near-field blocks and translation
 * schedules; alpha follows the creation parameters again (0,0: i/k).  robin
 * as for fmb_plan_create, at the new frequency.  The operator equals that of
 * fmb_plan_create at k with the same parameters.  On failure the plan holds
 * no operator: only fmb_plan_retune and fmb_plan_destroy accept it. */
int fmb_plan_retune(fmb_plan *pl, double k, const fmb_c64 *robin, void *stream);
/* Tile shape of the matrix-core translation kind (0 M2L at level, 1 M2M
 * level -> level-1, 2 L2L level-1 -> level) in the pass of precision prec:
 * mode (1: the batch in LDS, 2: in a global workspace, 3: in LDS on a
 * persistent grid, 0: VALU kernel), 16-column
 * tiles per batch nt (8 nt vectors), row tiles per unit mt, waves per
 * workgroup nw (4 or 8), bf = 1 where the rotations run in split-BF16, and
 * with autotune the timed milliseconds of one launch (else 0).  Any output
 * pointer may be NULL. */
int fmb_plan_tile_shape(const fmb_plan *pl, int kind, int level, int prec, int *mode,
                        int *nt, int *mt, int *nw, int *bf, double *ms);
/* Wavenumber of the operator the plan holds (0 after a failed retune). */
int fmb_plan_wavenumber(const fmb_plan *pl, double *k);
/* Device bytes held by the plan. */
int fmb_plan_memory(const fmb_plan *pl, double *bytes);

/* OprL (translation.cu:3412-3769), completed: y = A x for the Burton-Miller
 * operator (near field + P2M, M2M, M2L, L2L, L2P) on nrhs vectors.  On an
 * FMB_MIXED plan prec selects FMB_MIXED (FP32 pass) or FMB_FP64. */
int fmb_matvec(fmb_plan *pl, int prec, const fmb_c64 *x, fmb_c64 *y,
               int nrhs, void *stream);

/* The two halves of fmb_matvec: parts = 1 near field only, 2 far field only
 * (P2M .. L2P: the part the reference's OprL computes), 3 both. */
int fmb_matvec_parts(fmb_plan *pl, int prec, int parts, const fmb_c64 *x,
                     fmb_c64 *y, int nrhs, void *stream);

/* Pressure at field points off the surface from a surface solution x
 * ([nrhs][N] device, caller order): the exterior representation formula
 * p(y) = p_inc(y) + sum_j x_j int_j (dG/dn_y - sigma_j G) dS_y with the
 * plan's kernels, signs and panel rules (the close-panel rule when
 * |y - c_j| < near_ratio * radius_j).  pts: npts*3 host doubles; dirs:
 * nrhs*3 host doubles (unit plane waves added as the incident field) or
 * NULL for the scattered field; out: [nrhs][npts] device.  FP64, direct
 * O(N npts) summation.  The reference has none (its OprL stops at the
 * surface); it serves field-point objectives such as the sum of |p|^2
 * on a circle of points. */
int fmb_field_eval(const fmb_plan *pl, const fmb_c64 *x, int nrhs,
                   const double *pts, int npts, const double *dirs, fmb_c64 *out,
                   void *stream);

/* fmb_field_eval through the tree: an FP64 upward pass (P2M, M2M) on x,
 * then per point the multipole expansions of the boxes well separated from
 * it (their level-l box indices differ by 2 or more in some coordinate, the
 * separation of the M2L lists; points outside the bounding cube alike) and
 * the panels of the leaf boxes next to it by the direct rule.  The error is
 * the truncation error of the plan's orders; the cost per point is the
 * expansions it meets instead of all N panels.  Runs on the FP64 opset's
 * expansion buffers (as an FP64 fmb_matvec does) in passes of at most
 * nrhs_max vectors.  Non-finite point coordinates are refused
 * (FMB_ERR_ARG).  The reference has none. */
int fmb_field_eval_fmm(fmb_plan *pl, const fmb_c64 *x, int nrhs,
                       const double *pts, int npts, const double *dirs,
                       fmb_c64 *out, void *stream);

/* Right-hand side p_inc + alpha dp_inc/dn for unit plane waves e^{ik d.x}
 * (dirs: nrhs*3 host doubles).  The reference has none. */
int fmb_rhs_plane_wave(const fmb_plan *pl, const double *dirs, int nrhs,
                       fmb_c64 *b, void *stream);

typedef struct fmb_gmres_info {
  int iters;            /* inner Arnoldi steps (max over rhs)           */
  int matvecs_fp64;     /* FP64 FMM passes                              */
  int matvecs_work;     /* working-precision FMM passes                 */
  int restarts;
  double relres[64];    /* final true relative residual per rhs (<=64)  */
  double seconds;
} fmb_gmres_info;

/* GMRES(restart) to ||b - A x|| <= tol ||b|| (FP64 true residual), FP64
 * Arnoldi with classical Gram-Schmidt and reorthogonalisation.  On a
 * FMB_MIXED plan: iterative refinement with FP64 residuals around GMRES
 * on the FP32 pass, whose cycles stop at an estimated residual of tol / 2
 * (the FP64 check still applies tol).  x holds the initial guess.  The
 * reference has none. */
int fmb_gmres(fmb_plan *pl, const fmb_c64 *b, fmb_c64 *x, int nrhs,
              int restart, double tol, int maxit, fmb_gmres_info *info,
              void *stream);
/* ||b - A_outer x|| <= tol ||b|| (outer plan's FP64 pass) by iterative
 * refinement around GMRES on an inner plan: each cycle solves
 * A_inner d = r to inner_tol (0 < inner_tol < 1, relative to |r|; the
 * inner plan's work precision, stopped on the Arnoldi estimate) and sets
 * x += d.  The inner plan must share the outer plan's tree, wavenumber,
 * coupling and Robin coefficients (FMB_ERR_ARG otherwise); built with a
 * looser truncation (fmb_plan_params.eps) its passes are cheaper, and each
 * cycle contracts the residual by about inner_tol plus the relative
 * difference of the two operators.  maxit bounds the inner Arnoldi steps.
 * info: iters = inner Arnoldi steps, matvecs_fp64 = outer FP64 passes,
 * matvecs_work = inner plan passes (either precision), restarts = cycles,
 * relres = final outer true residuals.  The reference has no solver. */
int fmb_gmres_refined(fmb_plan *outer, fmb_plan *inner, const fmb_c64 *b, fmb_c64 *x,
                      int nrhs, int restart, double tol, double inner_tol, int maxit,
                      fmb_gmres_info *info, void *stream);
/* fmb_gmres_refined over several frequencies in lock step: outers[p] and
 * inners[p] as the pair of fmb_gmres_refined, columns of b and x as
 * fmb_gmres_batch (the inner solves run as one fmb_gmres_batch, graph as
 * there); a plan whose right-hand sides all meet tol takes no further outer
 * pass.  info: one per plan.  The reference solves one frequency at a
 * time. */
int fmb_gmres_batch_refined(fmb_plan *const *outers, fmb_plan *const *inners, const int *nrhs,
                            int nplans, int graph, const fmb_c64 *b, fmb_c64 *x, int restart,
                            double tol, double inner_tol, int maxit, fmb_gmres_info *info,
                            void *stream);
/* fmb_gmres_refined on A^T x = b (transposed passes of both plans; same
 * options, checks and info).  The reference has no adjoint. */
int fmb_gmres_refined_transpose(fmb_plan *outer, fmb_plan *inner, const fmb_c64 *b,
                                fmb_c64 *x, int nrhs, int restart, double tol,
                                double inner_tol, int maxit, fmb_gmres_info *info,
                                void *stream);
/* Initial guess for a sweep step from the solution x_from at wavenumber
 * k_from: x0_j = x_from_j e^{i (k - k_from) d_r . c_j} with c_j the centroid
 * of element j and d_r the unit plane-wave direction of right-hand side r
 * (dirs: nrhs*3 host doubles; NULL gives x0 = x_from), k the plan's
 * wavenumber.  It carries the incident phase across the step, so that
 * the illuminated part of the solution stays in place.  x_from and x0 may
 * alias.  The reference has no solver. */
int fmb_sweep_guess(const fmb_plan *pl, double k_from, const fmb_c64 *x_from,
                    const double *dirs, int nrhs, fmb_c64 *x0, void *stream);
/* GMRES over several plans at once (the reference has none): plans[p]
 * (one frequency each, any tree depth and orders; all on meshes with the
 * same element count) owns nrhs[p] right-hand sides; b and x are device
 * vectors [sum nrhs][N] with plan 0's columns first.  Every Arnoldi step
 * applies all plans in lock step (FP64 Arnoldi and true-residual
 * refinement as fmb_gmres, with the work precision shared by the plans).
 * graph = 1 captures the plans' FMM passes once per precision into one
 * hipGraph with one branch per plan (run concurrently, replayed every
 * step); a plan whose right-hand sides have all converged leaves the graph
 * at the next restart.  info: nplans reports (may be NULL).  A plan must not
 * be retuned or destroyed during the call. */
int fmb_gmres_batch(fmb_plan *const *plans, const int *nrhs, int nplans, int graph,
                    const fmb_c64 *b, fmb_c64 *x, int restart, double tol, int maxit,
                    fmb_gmres_info *info, void *stream);

/* y = A x for every plan of a batch (plans and columns as fmb_gmres_batch),
 * eagerly (graph = 0) or by replaying one captured hipGraph reps times (the
 * output of replay r in y + r * sum(nrhs) * N; eager: reps results of
 * repeated passes).  The reference has none. */
int fmb_matvec_batch(fmb_plan *const *plans, const int *nrhs, int nplans, int graph,
                     int prec, const fmb_c64 *x, fmb_c64 *y, int reps, void *stream);

/* y = A^T x (plain transpose, no conjugation) of the operator fmb_matvec
 * applies: the FMM pass run backwards with every operator transposed (L2P^T,
 * L2L^T, M2L^T, M2M^T, P2M^T; near-field blocks read by column).  parts as
 * for fmb_matvec_parts.  Needs rcr_path FMB_RCR_AUTO.  The reference has no
 * transposed operator; it serves adjoint solves. */
int fmb_matvec_transpose(fmb_plan *pl, int prec, int parts, const fmb_c64 *x,
                         fmb_c64 *y, int nrhs, void *stream);
/* fmb_gmres on A^T x = b (same options, refinement and info). */
int fmb_gmres_transpose(fmb_plan *pl, const fmb_c64 *b, fmb_c64 *x, int nrhs,
                        int restart, double tol, int maxit, fmb_gmres_info *info,
                        void *stream);

/* Adjoint gradients (the reference has none).  For the field-point objective
 *   J = sum_r sum_q |p_r(y_q)|^2,
 * with p_r the field of fmb_field_eval from the surface solution x ([nrhs][N]
 * device, A x = b solved beforehand for the plane waves dirs: nrhs*3 host
 * doubles, or NULL for the scattered field) at pts (npts*3 host doubles):
 *   J (host), may be NULL;
 *   dJ_dsigma[j] = dJ/dRe sigma_j + i dJ/dIm sigma_j ([N] device, caller
 *     order) for the plan's Robin coefficients (dp/dn = sigma p); for an
 *     admittance Y with sigma = i k Y, dJ/dY = -i k dJ/dsigma;
 *   dJ_dV[3v+a] = dJ/d(vertex v coordinate a) ([nvert*3] device) and
 *   dJ_dvn[v] = dJ_dV[v] . N_v along the unit area-weighted vertex normal
 *     N_v (fmb_tree_vertex_normals) ([nvert] device); these need dirs (the
 *     right-hand side moves with the collocation points).
 * Any of the three gradient outputs may be NULL, not all.  One adjoint
 * solve A^T lambda = F^T conj(p) (fmb_gmres_transpose with restart / tol /
 * maxit; its report in info, may be NULL).  Then
 *   dJ/dsigma_j = 2 conj(-x_j ((C^T lambda)_j + (Sf^T conj(p))_j)),
 * C = S + alpha K' + alpha I/2 (dA/dsigma_j's column factor), Sf the single
 * layer at the field points; and dJ = 2 Re [conj(p)^T dF x + lambda^T (db -
 * dA x)] per element, the element's own panel, collocation point, normal
 * and self term moved by central differences of its three vertices.  Needs
 * rcr_path FMB_RCR_AUTO. */
int fmb_adjoint_grad(fmb_plan *pl, const fmb_c64 *x, int nrhs, const double *dirs,
                     const double *pts, int npts, int restart, double tol, int maxit,
                     double *J, fmb_c64 *dJ_dsigma, double *dJ_dV, double *dJ_dvn,
                     fmb_gmres_info *info, void *stream);
/* fmb_adjoint_grad with the adjoint solve refined around GMRES on an inner
 * plan (fmb_gmres_refined_transpose with inner_tol; same checks of the
 * pair).  The reference has no adjoint. */
int fmb_adjoint_grad_refined(fmb_plan *pl, fmb_plan *inner, double inner_tol,
                             const fmb_c64 *x, int nrhs, const double *dirs, const double *pts,
                             int npts, int restart, double tol, int maxit, double *J,
                             fmb_c64 *dJ_dsigma, double *dJ_dV, double *dJ_dvn,
                             fmb_gmres_info *info, void *stream);
/* Precision of the Arnoldi matvecs of fmb_gmres on this plan (FMB_MIXED only
 * on a plan created with FMB_MIXED).  The reference has no solver. */
int fmb_plan_set_work_precision(fmb_plan *pl, int prec);
void fmb_plan_destroy(fmb_plan *pl);

/* Per-phase timing of the last fmb_matvec (ms): near, p2m, m2m, m2l, l2l,
 * l2p, total; and the M2L kernel's event time per level. */
int fmb_last_timing(const fmb_plan *pl, double *phase7, double *m2l16);
/* Enables HIP-event timing of each phase inside fmb_matvec. */
int fmb_set_timing(fmb_plan *pl, int on);

#ifdef __cplusplus
}
#endif
#endif
