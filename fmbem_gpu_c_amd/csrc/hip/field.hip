// field.hip -- pressure at field points off the surface from a surface
// solution, by the exterior representation formula
//   p(y) = p_inc(y) + sum_j x_j int_j (dG/dn_y - sigma_j G) dS_y
// with the same kernels, signs and panel rules as the collocation entries
// (near.hip): the refined rule when |y - c_j| < ratio * radius_j.  The
// reference has no field evaluation.  One workgroup per field point reduces
// over all panels in FP64; R right-hand sides per pass.
//
// Its transpose, for adjoint objectives: given weights u[r][q] at the field
// points, per element j (caller order)
//   kt[r][j] = sum_q u[r][q] int_j (dG/dn_y - sigma_j G)(y_q, y) dS_y
//   st[r][j] = sum_q u[r][q] int_j G(y_q, y) dS_y
// one thread per element, R right-hand sides per pass.
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

__global__ void __launch_bounds__(kT)
    k_field_t(const double *__restrict__ pts, int npts, const double *__restrict__ tri,
              const double *__restrict__ geo, const int *__restrict__ perm, int N,
              double k, const double2 *__restrict__ robin,
              const double *__restrict__ quad, int nq, const double *__restrict__ quadn,
              int nqn, double ratio, const double2 *__restrict__ u, int nrhs, int r0,
              double2 *__restrict__ kt, double2 *__restrict__ st) {
  const int j = blockIdx.x * kT + threadIdx.x;
  if (j >= N) return;
  const int nr = min(kR, nrhs - r0);
  const double *gj = geo + 8 * (size_t)j;
  const double nz[3] = {0.0, 0.0, 0.0};
  const double2 sg = robin ? robin[j] : make_double2(0.0, 0.0);
  double2 ak[kR], as[kR];
#pragma unroll
  for (int r = 0; r < kR; r++) ak[r] = as[r] = make_double2(0.0, 0.0);
  for (int q = 0; q < npts; q++) {
    const double y[3] = {pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]};
    const double dx = y[0] - gj[0], dy = y[1] - gj[1], dz = y[2] - gj[2];
    const bool close = dx * dx + dy * dy + dz * dz < ratio * ratio * gj[7] * gj[7];
    double2 o[4];
    panel_sum(k, y, nz, tri + 9 * (size_t)j, gj + 3, gj[6], close ? quadn : quad,
              close ? nqn : nq, o);
    const double2 g = make_double2(o[1].x - (sg.x * o[0].x - sg.y * o[0].y),
                                   o[1].y - (sg.x * o[0].y + sg.y * o[0].x));
#pragma unroll
    for (int r = 0; r < kR; r++) {
      if (r < nr) {
        const double2 w = u[(size_t)(r0 + r) * npts + q];
        ak[r].x += g.x * w.x - g.y * w.y;
        ak[r].y += g.x * w.y + g.y * w.x;
        as[r].x += o[0].x * w.x - o[0].y * w.y;
        as[r].y += o[0].x * w.y + o[0].y * w.x;
      }
    }
  }
  const int c = perm[j];
#pragma unroll
  for (int r = 0; r < kR; r++) {
    if (r < nr) {
      kt[(size_t)(r0 + r) * N + c] = ak[r];
      st[(size_t)(r0 + r) * N + c] = as[r];
    }
  }
}

}  // namespace

extern "C" int fmbh_field_t(const fmb_plan *pl, const fmb_c64 *u, int nrhs,
                            const double *d_pts, int npts, fmb_c64 *kt, fmb_c64 *st,
                            void *stream) {
  const fmb_tree *t = pl->tree;
  hipStream_t s = (hipStream_t)stream;
  const int N = t->nelem;
  for (int r0 = 0; r0 < nrhs; r0 += kR) {
    hipLaunchKernelGGL(k_field_t, dim3((N + kT - 1) / kT), dim3(kT), 0, s, d_pts, npts,
                       t->d_tri, t->d_geo, t->d_perm, N, pl->k,
                       (const double2 *)pl->d_robin, pl->d_quad, pl->nq, pl->d_quadn,
                       pl->nqn, pl->prm.near_ratio, (const double2 *)u, nrhs, r0,
                       (double2 *)kt, (double2 *)st);
    FMB_LAUNCH_CHECK();
  }
  return FMB_OK;
}

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
