// field.hip -- pressure at field points off the surface from a surface
// solution, by the exterior representation formula
//   p(y) = p_inc(y) + sum_j x_j int_j (dG/dn_y - sigma_j G) dS_y
// with the same kernels, signs and panel rules as the collocation entries
// (near.hip): the refined rule when |y - c_j| < ratio * radius_j.  The
// reference has no field evaluation.  One workgroup per field point reduces
// over all panels in FP64; R right-hand sides per pass.
#include "fmb_device.h"

namespace {

constexpr int kT = 256, kR = 8;

__global__ void __launch_bounds__(kT)
    k_field(const double *__restrict__ pts, int npts, const double *__restrict__ tri,
            const double *__restrict__ geo, const int *__restrict__ perm, int N,
            double k, const double2 *__restrict__ robin,
            const double *__restrict__ quad, int nq, const double *__restrict__ quadn,
            int nqn, double ratio, const double2 *__restrict__ x, int nrhs, int r0,
            const double *__restrict__ dirs, double2 *__restrict__ out) {
  __shared__ double2 red[kT / 64][kR];
  const int q = blockIdx.x;
  const double y[3] = {pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]};
  const double nz[3] = {0.0, 0.0, 0.0};
  const int nr = min(kR, nrhs - r0);
  double2 acc[kR];
#pragma unroll
  for (int r = 0; r < kR; r++) acc[r] = make_double2(0.0, 0.0);
  for (int j = threadIdx.x; j < N; j += kT) {
    const double *gj = geo + 8 * (size_t)j;
    const double dx = y[0] - gj[0], dy = y[1] - gj[1], dz = y[2] - gj[2];
    const bool close = dx * dx + dy * dy + dz * dz < ratio * ratio * gj[7] * gj[7];
    double2 o[4];
    panel_sum(k, y, nz, tri + 9 * (size_t)j, gj + 3, gj[6], close ? quadn : quad,
              close ? nqn : nq, o);
    double2 g = o[1];
    if (robin) {
      const double2 s = robin[j];
      g.x -= s.x * o[0].x - s.y * o[0].y;
      g.y -= s.x * o[0].y + s.y * o[0].x;
    }
    const int c = perm[j];
#pragma unroll
    for (int r = 0; r < kR; r++) {
      if (r < nr) {
        const double2 v = x[(size_t)(r0 + r) * N + c];
        acc[r].x += g.x * v.x - g.y * v.y;
        acc[r].y += g.x * v.y + g.y * v.x;
      }
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < kR; r++) {
    double re = acc[r].x, im = acc[r].y;
    for (int o = 32; o > 0; o >>= 1) {
      re += __shfl_down(re, o, 64);
      im += __shfl_down(im, o, 64);
    }
    if (lane == 0) red[w][r] = make_double2(re, im);
  }
  __syncthreads();
  if (threadIdx.x < nr) {
    const int r = threadIdx.x;
    double2 s = make_double2(0.0, 0.0);
    for (int i = 0; i < kT / 64; i++) {
      s.x += red[i][r].x;
      s.y += red[i][r].y;
    }
    if (dirs) {
      const double *d = dirs + 3 * (r0 + r);
      double sn, cs;
      sincos(k * (d[0] * y[0] + d[1] * y[1] + d[2] * y[2]), &sn, &cs);
      s.x += cs;
      s.y += sn;
    }
    out[(size_t)(r0 + r) * npts + q] = s;
  }
}

}  // namespace

extern "C" int fmbh_field(const fmb_plan *pl, const fmb_c64 *x, int nrhs,
                          const double *d_pts, int npts, const double *d_dirs,
                          fmb_c64 *out, void *stream) {
  const fmb_tree *t = pl->tree;
  hipStream_t st = (hipStream_t)stream;
  for (int r0 = 0; r0 < nrhs; r0 += kR) {
    hipLaunchKernelGGL(k_field, dim3(npts), dim3(kT), 0, st, d_pts, npts, t->d_tri,
                       t->d_geo, t->d_perm, t->nelem, pl->k,
                       (const double2 *)pl->d_robin, pl->d_quad, pl->nq, pl->d_quadn,
                       pl->nqn, pl->prm.near_ratio, (const double2 *)x, nrhs, r0,
                       d_dirs, (double2 *)out);
    FMB_LAUNCH_CHECK();
  }
  return FMB_OK;
}
