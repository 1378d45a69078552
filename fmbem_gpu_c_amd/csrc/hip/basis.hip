// basis.hip -- per-element P2M and L2P bases, evaluated once per plan in FP64
// and stored scaled in the compute precision.
//
//   P2M_j[n,m] = ik sum_q w_q A_j [ -dR_n^{-m}/dn_j + sigma_j R_n^{-m} ](y_q - c)
//                / w_n                    (bottomLevelCoeff_L, integral.cu:821-836,
//                                          triElemIntegral_R / _pRpn, 251-365)
//   L2P_i[n,m] = [ R_n^m + alpha dR_n^m/dn_i ](x_i - c) / w_n
// c is the leaf-box centre; dR/dn follows dirDev_R (integral.cu:284-331).
// With dsig set the kernel writes dP2M_j/dsigma_j = ik sum_q w_q A_j R_n^{-m}
// / w_n instead (and no L2P), the column factor of the sigma derivative.
// One wavefront per element: lane 0 runs the serial Bessel/Legendre
// diagonals, the lanes share the Legendre columns and the R table in LDS.
#include "fmb_device.h"

namespace {

constexpr int WG = 64;

template <typename T>
__global__ void __launch_bounds__(WG)
    k_bases(int N, int pL, double k, double alre, double alim,
            const double *__restrict__ tri, const double *__restrict__ geo,
            const int *__restrict__ ebox, const int *__restrict__ keys, int L,
            double D, double p0, double p1, double p2,
            const double2 *__restrict__ robin, const double *__restrict__ quad,
            int nq, const double *__restrict__ wsc, T *__restrict__ P2M,
            T *__restrict__ L2P, int dsig) {
  extern __shared__ double sm[];
  int nm = pL;  // R needed up to degree pL for the derivative of degree pL-1
  int ntab = (nm + 1) * (nm + 2) / 2;
  const int ps = pL * pL;
  double2 *ph = (double2 *)sm;
  double2 *Rt = ph + (nm + 1);
  double2 *acc = Rt + (nm + 1) * (nm + 1);
  double *jn = (double *)(acc + ps);
  double *diag = jn + (nm + 1);
  double *Pt = diag + (nm + 1);
  double *g3 = Pt + ntab;
  int j = blockIdx.x;
  int lane = threadIdx.x;
  // leaf-box centre
  int key = keys[ebox[j]];
  int tx = 0, ty = 0, tz = 0;
  for (int i = 0; i < L; i++) {
    int d = (key >> (3 * i)) & 7;
    tx |= ((d >> 2) & 1) << i;
    ty |= ((d >> 1) & 1) << i;
    tz |= (d & 1) << i;
  }
  double h = ldexp(1.0, -L);
  double c[3] = {D * (h * tx + 0.5 * h) + p0, D * (h * ty + 0.5 * h) + p1,
                 D * (h * tz + 0.5 * h) + p2};
  const double *tr = tri + 9 * (size_t)j;
  const double *g = geo + 8 * (size_t)j;
  double nr[3] = {g[3], g[4], g[5]};
  double area = g[6];
  double2 sg = robin ? robin[j] : make_double2(0.0, 0.0);
  for (int f = lane; f < ps; f += WG) acc[f] = make_double2(0.0, 0.0);
  for (int q = 0; q < nq; q++) {
    double v[3];
    const double *qq = quad + 4 * q;
    for (int a = 0; a < 3; a++)
      v[a] = qq[0] * tr[a] + qq[1] * tr[3 + a] + qq[2] * tr[6 + a] - c[a];
    double w = qq[3] * area;
    r_table(nm, k, v, jn, diag, Pt, ph, Rt, g3);
    for (int f = lane; f < ps; f += WG) {
      int n = (int)sqrt((double)f);
      while (n * n > f) n--;
      while ((n + 1) * (n + 1) <= f) n++;
      int m = f - n * n - n;
      double2 R = Rt[n * n + n - m];
      double2 dR = dRdn(Rt, n, -m, k, nr);
      // val = -dR + sigma R (dsig: R) ; P2M += i k w val
      double vr = dsig ? R.x : -dR.x + sg.x * R.x - sg.y * R.y;
      double vi = dsig ? R.y : -dR.y + sg.x * R.y + sg.y * R.x;
      acc[f].x += -k * w * vi;
      acc[f].y += k * w * vr;
    }
    __syncthreads();
  }
  for (int f = lane; f < ps; f += WG) {
    int n = (int)sqrt((double)f);
    while (n * n > f) n--;
    while ((n + 1) * (n + 1) <= f) n++;
    double s = 1.0 / wsc[n];
    P2M[2 * ((size_t)j * ps + f)] = (T)(acc[f].x * s);
    P2M[2 * ((size_t)j * ps + f) + 1] = (T)(acc[f].y * s);
  }
  if (!L2P) return;
  // L2P at the centroid
  double v[3] = {g[0] - c[0], g[1] - c[1], g[2] - c[2]};
  r_table(nm, k, v, jn, diag, Pt, ph, Rt, g3);
  for (int f = lane; f < ps; f += WG) {
    int n = (int)sqrt((double)f);
    while (n * n > f) n--;
    while ((n + 1) * (n + 1) <= f) n++;
    int m = f - n * n - n;
    double2 R = Rt[f];
    double2 dR = dRdn(Rt, n, m, k, nr);
    double er = R.x + alre * dR.x - alim * dR.y;
    double ei = R.y + alre * dR.y + alim * dR.x;
    double s = 1.0 / wsc[n];
    L2P[2 * ((size_t)j * ps + f)] = (T)(er * s);
    L2P[2 * ((size_t)j * ps + f) + 1] = (T)(ei * s);
  }
}

}  // namespace

static int launch_bases(fmb_plan *pl, int op, const fmb_c64 *robin, void *P2M,
                        void *L2P, int dsig, void *stream) {
  const fmb_tree *t = pl->tree;
  int L = t->lmax, pL = pl->p[L];
  int nm = pL;
  size_t lds = sizeof(double) * (2 * (nm + 1) + (nm + 1) * (nm + 2) / 2 + 4) +
               sizeof(double2) * ((nm + 1) + (nm + 1) * (nm + 1) + pL * pL);
  if (lds > 160 * 1024)
    return fmb_set_error(FMB_ERR_RANGE, "leaf order %d too large for bases",
                         pL);
  hipStream_t st = (hipStream_t)stream;
  fmb_opset *o = &pl->ops[op];
  if (o->prec == FMB_FP64) {
    auto kern = k_bases<double>;
    if (lds > 64 * 1024)
      FMB_HIP_CHECK(hipFuncSetAttribute(
          (const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
          160 * 1024));
    hipLaunchKernelGGL(kern, dim3(t->nelem), dim3(WG), lds, st, t->nelem, pL,
                       pl->k, pl->alpha_re, pl->alpha_im, t->d_tri, t->d_geo,
                       t->d_elem_box, t->d_keys[L], L, t->D, t->pmin[0],
                       t->pmin[1], t->pmin[2], (const double2 *)robin,
                       pl->d_quad, pl->nq, pl->d_scale[L], (double *)P2M,
                       (double *)L2P, dsig);
  } else {
    auto kern = k_bases<float>;
    if (lds > 64 * 1024)
      FMB_HIP_CHECK(hipFuncSetAttribute(
          (const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
          160 * 1024));
    hipLaunchKernelGGL(kern, dim3(t->nelem), dim3(WG), lds, st, t->nelem, pL,
                       pl->k, pl->alpha_re, pl->alpha_im, t->d_tri, t->d_geo,
                       t->d_elem_box, t->d_keys[L], L, t->D, t->pmin[0],
                       t->pmin[1], t->pmin[2], (const double2 *)robin,
                       pl->d_quad, pl->nq, pl->d_scale[L], (float *)P2M,
                       (float *)L2P, dsig);
  }
  FMB_LAUNCH_CHECK();
  FMB_HIP_CHECK(hipStreamSynchronize(st));
  return FMB_OK;
}

extern "C" int fmbh_bases(fmb_plan *pl, int op, const fmb_c64 *robin,
                          void *stream) {
  return launch_bases(pl, op, robin, pl->ops[op].d_P2M, pl->ops[op].d_L2P, 0, stream);
}

extern "C" int fmbh_bases_dsigma(fmb_plan *pl, int op, void *out, void *stream) {
  return launch_bases(pl, op, NULL, out, NULL, 1, stream);
}
