/* host.h -- helpers shared by the C host modules. */
#ifndef FMB_HOST_H
#define FMB_HOST_H
#include "../fmb_internal.h"
void fmbx_gauss_legendre(int n, double *x, double *w);
int fmbx_tri_rule(int rule, int order, double **out);
void fmbx_abs_hn(int nmax, double x, double *out);
/* matrix-core translations (rcr.c): phase tables and padded rotation blocks
 * of every opset (rotation tier), then schedules, padded coaxial blocks and
 * workspaces (frequency tier) */
int fmbx_rcr_rot_setup(fmb_plan *pl, void *stream);
int fmbx_rcr_setup(fmb_plan *pl, void *stream);
void fmbx_rcr_free(fmb_plan *pl);
int fmbx_rcr_run(const fmb_plan *pl, int op, int kind, int l, int nrhs,
                 void *stream);
/* transposed translations, coefficients kept multiplied by (-1)^n */
int fmbx_rcr_run_t(const fmb_plan *pl, int op, int kind, int l, int nrhs,
                   void *stream);
/* transposed FMM pass (plan.c); near / basis NULL: the opset's own */
int fmbx_matvec_t(fmb_plan *pl, int op, int parts, const void *near,
                  const void *basis, const fmb_c64 *x, fmb_c64 *y, int nrhs,
                  void *stream);
#endif
