/* ops.c -- standalone operator generation in the reference layouts:
 * genSparseRotMats (translation.cu:1297-1394) and
 * gen{RR,SR}SparseCoaxTransMat (translation.cu:822-903, 982-1062), built on
 * the device in FP64 and copied to the caller's host buffer. */
#include <stdlib.h>

#include "../fmb_internal.h"
#include "../hip/fmb_hip.h"

int fmb_gen_rot_mats(const double *angles, int n, int p, fmb_c64 *out) {
  if (!angles || n < 1 || p < 1 || !out)
    return fmb_set_error(FMB_ERR_ARG, "fmb_gen_rot_mats: invalid arguments");
  int rc = fmbh_init();
  if (rc) return rc;
  size_t rs = (size_t)fmb_rot_size(p);
  double *betas = (double *)malloc(sizeof(double) * n);
  if (!betas) return fmb_set_error(FMB_ERR_NOMEM, "fmb_gen_rot_mats");
  for (int i = 0; i < n; i++) betas[i] = angles[3 * i + 1];
  double *d_H = NULL;
  fmb_c64 *d_out = NULL;
  if ((rc = fmbh_malloc((void **)&d_H, sizeof(double) * rs * n))) goto done;
  if ((rc = fmbh_malloc((void **)&d_out, sizeof(fmb_c64) * rs * n))) goto done;
  if ((rc = fmbh_gen_H(betas, n, p, d_H, NULL))) goto done;
  if ((rc = fmbh_rot_phase(d_H, n, p, angles, d_out, NULL))) goto done;
  rc = fmbh_d2h(out, d_out, sizeof(fmb_c64) * rs * n, NULL);
done:
  free(betas);
  fmbh_free(d_H);
  fmbh_free(d_out);
  return rc;
}

int fmb_gen_coax_mats(int kind, double k, const double *t, int n, int p,
                      fmb_c64 *out) {
  if ((kind != FMB_COAX_RR && kind != FMB_COAX_SR) || !t || n < 1 || p < 1 ||
      !out)
    return fmb_set_error(FMB_ERR_ARG, "fmb_gen_coax_mats: invalid arguments");
  int rc = fmbh_init();
  if (rc) return rc;
  size_t cs = (size_t)fmb_coax_size(p);
  fmb_c64 *d_out = NULL;
  if ((rc = fmbh_malloc((void **)&d_out, sizeof(fmb_c64) * cs * n))) return rc;
  if (!(rc = fmbh_gen_coax(kind, k, t, n, p, d_out, NULL)))
    rc = fmbh_d2h(out, d_out, sizeof(fmb_c64) * cs * n, NULL);
  fmbh_free(d_out);
  return rc;
}
