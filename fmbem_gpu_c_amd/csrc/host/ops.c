/* ops.c -- standalone operator generation in the reference layouts:
 * genSparseRotMats (translation.cu:1297-1394) and
 * gen{RR,SR}SparseCoaxTransMat (translation.cu:822-903, 982-1062), built on
 * the device in FP64 and copied to the caller's host buffer; and their
 * truncation to a lower order (reduceSparseRotMat / reduceSparseCoaxTransMat,
 * translation.cu:22-56) on the host. */
#include <stdlib.h>
#include <string.h>

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

/* fmb_gen_coax_mats without allocations: device input and output, the
 * recurrence table in the library's fixed device scratch (no counterpart in
 * the reference). */
int fmb_gen_coax_mats_noalloc(int kind, double k, const double *d_t, int n, int p,
                              fmb_c64 *d_out, void *stream) {
  if ((kind != FMB_COAX_RR && kind != FMB_COAX_SR) || !d_t || n < 1 || p < 1 || !d_out)
    return fmb_set_error(FMB_ERR_ARG, "fmb_gen_coax_mats_noalloc: invalid arguments");
  return fmbh_gen_coax_static(kind, k, d_t, n, p, d_out, stream);
}

size_t fmb_coax_scratch_bytes(void) { return fmbh_coax_scratch_bytes(); }

/* reduceSparseRotMat (translation.cu:22-26): the n-blocks n < p2 of a
 * rotation matrix of order p1 are its first fmb_rot_size(p2) entries. */
int fmb_reduce_rot_mat(const fmb_c64 *in, int p1, fmb_c64 *out, int p2) {
  if (!in || !out || p2 < 1 || p2 > p1)
    return fmb_set_error(FMB_ERR_ARG, "fmb_reduce_rot_mat: orders %d -> %d", p1, p2);
  memmove(out, in, sizeof(fmb_c64) * (size_t)fmb_rot_size(p2));
  return FMB_OK;
}

/* offset of the |m|-block of a coaxial matrix of order p (sparseCoaxMatBlockIdx,
 * translation.cu:28-32): blocks (p - m)^2 for m = 0, 1, ... */
static size_t coax_block(int p, int m) {
  return (size_t)m * (2 * (size_t)m * m - (3 + 6 * (size_t)p) * m + 1 + 6 * ((size_t)p + (size_t)p * p)) / 6;
}

/* reduceSparseCoaxTransMat (translation.cu:34-56): block m < p2 of the
 * order-p2 matrix is the leading (p2-m) x (p2-m) part of block m of the
 * order-p1 matrix, both column-major. */
int fmb_reduce_coax_mat(const fmb_c64 *in, int p1, fmb_c64 *out, int p2) {
  if (!in || !out || p2 < 1 || p2 > p1)
    return fmb_set_error(FMB_ERR_ARG, "fmb_reduce_coax_mat: orders %d -> %d", p1, p2);
  if ((const void *)in == (const void *)out && p1 != p2)
    return fmb_set_error(FMB_ERR_ARG, "fmb_reduce_coax_mat: out must not alias in");
  for (int m = 0; m < p2; m++) {
    const fmb_c64 *a = in + coax_block(p1, m);
    fmb_c64 *b = out + coax_block(p2, m);
    const int l1 = p1 - m, l2 = p2 - m;
    for (int c = 0; c < l2; c++)
      for (int r = 0; r < l2; r++) b[(size_t)c * l2 + r] = a[(size_t)c * l1 + r];
  }
  return FMB_OK;
}
