// shape.hip -- shape sensitivities of a field-point objective for the
// adjoint (the reference has no adjoint).
//
// With x (A x = b), lambda (A^T lambda = F^T u) and u = conj(p) the change
// of J = sum |p|^2 under a change of the geometry is
//   dJ = 2 Re [ u^T dF x + lambda^T (db - dA x) ].
// Each element e owns the terms in which its geometry g_e (three vertices)
// enters: the field panel F_qe, the right-hand side b_e (collocation point
// and normal), row e of A (collocation at c_e, n_e), column e of A (panel
// integrals over e) and the self term A_ee.  Per element
//   Q_e(g) = sum_r [ x_e sum_q u_q F_qe(g) + lam_e b_e(g) - lam_e R_e(g)
//                    - C_e(g) x_e + lam_e A_ee(g) x_e ],
//   R_e = sum_j A_ej x_j (near row + L2P_e(g) . L_box),
//   C_e = sum_i lam_i A_ie (near column + P2M_e(g) . M'_box),
// where L_box is the forward pass's leaf local expansion of x and M'_box
// the transposed pass's leaf multipole expansion of lambda; the diagonal
// enters R and C once each and is added back once.  dJ/dV = 2 Re sum_e
// dQ_e/dV; the partial derivatives of Q_e in the nine vertex coordinates of
// e are central differences (step hrel * element radius) of Q_e evaluated
// with exactly the kernels, rules and tables of the plan, so they are the
// derivatives of the FMM operator the plan applies.
#include "fmb_device.h"

namespace {

constexpr int W = 64, kR = 8;

struct ShapeArgs {
  int N, nrhs, r0, npts, L, pL;
  double k, alre, alim, D, p0, p1, p2, ratio, hrel;
  const double *tri, *geo;
  const int *ebox, *keys, *bstart, *nstart, *nbox;
  const double2 *robin;
  const double *quad, *quadn, *edge;
  int nq, nqn, nedge;
  const double *wsc;
  const double2 *xs, *lam;  // sorted order [nrhs][N]
  const double2 *Lx, *Mt;   // [box][nrhs][pL^2]
  const double *pts;
  const double2 *u;         // [nrhs][npts]
  const double *dirs;       // [nrhs][3]
  double2 *Q;               // [N][18], accumulated
};

__device__ __forceinline__ double2 cmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd2(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 wsum2(double2 v) {
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_down(v.x, o, 64);
    v.y += __shfl_down(v.y, o, 64);
  }
  return v;
}

// A_ij of the Burton-Miller operator (i != j) from the panel integrals
__device__ __forceinline__ double2 entry(const double2 o[4], double2 al, double2 sj) {
  double2 aT = cmul2(al, o[3]), aKp = cmul2(al, o[2]);
  double2 st = cmul2(sj, make_double2(o[0].x + aKp.x, o[0].y + aKp.y));
  return make_double2(-o[1].x - aT.x + st.x, -o[1].y - aT.y + st.y);
}

// one 64-lane workgroup per (element e, perturbation s)
__global__ void __launch_bounds__(W) k_shape_q(ShapeArgs a) {
  extern __shared__ double sm[];
  __shared__ double G[9], geo_g[8], ctr[3];
  __shared__ int cpre[28], cbox[27], ncb;
  __shared__ double2 aee_s;
  const int e = blockIdx.x, s = blockIdx.y, lane = threadIdx.x;
  const int nr = min(kR, a.nrhs - a.r0);
  const int nm = a.pL, ps = a.pL * a.pL;
  const int ntab = (nm + 1) * (nm + 2) / 2;
  double2 *ph = (double2 *)sm;
  double2 *Rt = ph + (nm + 1);
  double2 *acc = Rt + (nm + 1) * (nm + 1);
  double *jn = (double *)(acc + ps);
  double *diag = jn + (nm + 1);
  double *Pt = diag + (nm + 1);
  double *g3 = Pt + ntab;
  const int box = a.ebox[e];
  if (lane == 0) {
    const double *t0 = a.tri + 9 * (size_t)e;
    for (int i = 0; i < 9; i++) G[i] = t0[i];
    const int pi = s >> 1, vs = pi / 3, ax = pi % 3;
    const double h = a.hrel * a.geo[8 * (size_t)e + 7];
    G[3 * vs + ax] += (s & 1) ? -h : h;
    // geometry exactly as the tree derives it
    double c[3], u1[3], u2[3];
    for (int q = 0; q < 3; q++) {
      c[q] = G[q] + 2.0 / 3.0 * (0.5 * (G[3 + q] + G[6 + q]) - G[q]);
      u1[q] = G[q] - G[6 + q];
      u2[q] = G[3 + q] - G[6 + q];
    }
    double nrm[3] = {u1[1] * u2[2] - u1[2] * u2[1], u1[2] * u2[0] - u1[0] * u2[2],
                     u1[0] * u2[1] - u1[1] * u2[0]};
    double J = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    double rad = 0.0;
    for (int v = 0; v < 3; v++) {
      double d2 = 0.0;
      for (int q = 0; q < 3; q++) d2 += (G[3 * v + q] - c[q]) * (G[3 * v + q] - c[q]);
      rad = fmax(rad, d2);
    }
    for (int q = 0; q < 3; q++) {
      geo_g[q] = c[q];
      geo_g[3 + q] = nrm[q] / J;
    }
    geo_g[6] = 0.5 * J;
    geo_g[7] = sqrt(rad);
    // leaf-box centre
    int key = a.keys[box], tx = 0, ty = 0, tz = 0;
    for (int i = 0; i < a.L; i++) {
      int d = (key >> (3 * i)) & 7;
      tx |= ((d >> 2) & 1) << i;
      ty |= ((d >> 1) & 1) << i;
      tz |= (d & 1) << i;
    }
    double hb = ldexp(1.0, -a.L);
    ctr[0] = a.D * (hb * tx + 0.5 * hb) + a.p0;
    ctr[1] = a.D * (hb * ty + 0.5 * hb) + a.p1;
    ctr[2] = a.D * (hb * tz + 0.5 * hb) + a.p2;
    // near neighbourhood of the box
    int acc0 = 0, q0 = a.nstart[box], q1 = a.nstart[box + 1];
    ncb = q1 - q0;
    for (int q = q0; q < q1; q++) {
      cpre[q - q0] = acc0;
      cbox[q - q0] = a.nbox[q];
      acc0 += a.bstart[a.nbox[q] + 1] - a.bstart[a.nbox[q]];
    }
    cpre[q1 - q0] = acc0;
    aee_s = make_double2(0.0, 0.0);
  }
  __syncthreads();
  const double c[3] = {geo_g[0], geo_g[1], geo_g[2]};
  const double nrm[3] = {geo_g[3], geo_g[4], geo_g[5]};
  const double area = geo_g[6], rad = geo_g[7];
  const double2 al = make_double2(a.alre, a.alim);
  const double2 se = a.robin ? a.robin[e] : make_double2(0.0, 0.0);
  const size_t N = a.N;
  double2 fld[kR], row[kR], col[kR];
#pragma unroll
  for (int r = 0; r < kR; r++) fld[r] = row[r] = col[r] = make_double2(0.0, 0.0);
  // field panel over e at the objective points
  const double nz[3] = {0.0, 0.0, 0.0};
  for (int q = lane; q < a.npts; q += W) {
    const double *y = a.pts + 3 * q;
    const double dx = y[0] - c[0], dy = y[1] - c[1], dz = y[2] - c[2];
    const bool close = dx * dx + dy * dy + dz * dz < a.ratio * a.ratio * rad * rad;
    double2 o[4];
    panel_sum(a.k, y, nz, G, nrm, area, close ? a.quadn : a.quad, close ? a.nqn : a.nq, o);
    const double2 f = make_double2(o[1].x - (se.x * o[0].x - se.y * o[0].y),
                                   o[1].y - (se.x * o[0].y + se.y * o[0].x));
#pragma unroll
    for (int r = 0; r < kR; r++)
      if (r < nr) fld[r] = cadd2(fld[r], cmul2(a.u[(size_t)(a.r0 + r) * a.npts + q], f));
  }
  // near row (collocation at c, n) and column (panel e) over the neighbourhood
  const int ncol = cpre[ncb];
  for (int uu = lane; uu < ncol; uu += W) {
    int q = 0;
    while (q + 1 < ncb && cpre[q + 1] <= uu) q++;
    const int j = a.bstart[cbox[q]] + (uu - cpre[q]);
    if (j == e) {
      double2 S, T;
      self_terms(a.k, G, c, a.edge, a.nedge, S, T);
      const double2 aT = cmul2(al, T), sS = cmul2(se, S), as = cmul2(al, se);
      const double2 v = make_double2(0.5 + 0.5 * as.x - aT.x + sS.x,
                                     0.5 * as.y - aT.y + sS.y);
      aee_s = v;
#pragma unroll
      for (int r = 0; r < kR; r++)
        if (r < nr) {
          row[r] = cadd2(row[r], cmul2(v, a.xs[(size_t)(a.r0 + r) * N + e]));
          col[r] = cadd2(col[r], cmul2(v, a.lam[(size_t)(a.r0 + r) * N + e]));
        }
      continue;
    }
    const double *gj = a.geo + 8 * (size_t)j;
    const double2 sj = a.robin ? a.robin[j] : make_double2(0.0, 0.0);
    const double dx = c[0] - gj[0], dy = c[1] - gj[1], dz = c[2] - gj[2];
    const double d2 = dx * dx + dy * dy + dz * dz;
    double2 o[4];
    // A_ej: collocation at the perturbed centroid, panel j
    bool close = d2 < a.ratio * a.ratio * gj[7] * gj[7];
    panel_sum(a.k, c, nrm, a.tri + 9 * (size_t)j, gj + 3, gj[6], close ? a.quadn : a.quad,
              close ? a.nqn : a.nq, o);
    const double2 aej = entry(o, al, sj);
    // A_je: collocation at c_j, the perturbed panel e
    close = d2 < a.ratio * a.ratio * rad * rad;
    panel_sum(a.k, gj, gj + 3, G, nrm, area, close ? a.quadn : a.quad,
              close ? a.nqn : a.nq, o);
    const double2 aje = entry(o, al, se);
#pragma unroll
    for (int r = 0; r < kR; r++)
      if (r < nr) {
        row[r] = cadd2(row[r], cmul2(aej, a.xs[(size_t)(a.r0 + r) * N + j]));
        col[r] = cadd2(col[r], cmul2(aje, a.lam[(size_t)(a.r0 + r) * N + j]));
      }
  }
  // far row: L2P basis of the perturbed collocation point . L_box
  {
    double v[3] = {c[0] - ctr[0], c[1] - ctr[1], c[2] - ctr[2]};
    r_table(nm, a.k, v, jn, diag, Pt, ph, Rt, g3);
    for (int f = lane; f < ps; f += W) {
      const int n = isqrt_fix(f), m = f - n * n - n;
      const double2 R = Rt[f], dR = dRdn(Rt, n, m, a.k, nrm);
      const double is = 1.0 / a.wsc[n];
      const double2 E = make_double2((R.x + a.alre * dR.x - a.alim * dR.y) * is,
                                     (R.y + a.alre * dR.y + a.alim * dR.x) * is);
#pragma unroll
      for (int r = 0; r < kR; r++)
        if (r < nr)
          row[r] = cadd2(row[r], cmul2(E, a.Lx[((size_t)box * a.nrhs + a.r0 + r) * ps + f]));
    }
  }
  // far column: P2M basis of the perturbed panel . M'_box ((-1)^n: the
  // transposed pass's coefficients carry that factor)
  for (int f = lane; f < ps; f += W) acc[f] = make_double2(0.0, 0.0);
  for (int q = 0; q < a.nq; q++) {
    const double *qq = a.quad + 4 * q;
    double v[3];
    for (int i = 0; i < 3; i++)
      v[i] = qq[0] * G[i] + qq[1] * G[3 + i] + qq[2] * G[6 + i] - ctr[i];
    const double w = qq[3] * area;
    r_table(nm, a.k, v, jn, diag, Pt, ph, Rt, g3);
    for (int f = lane; f < ps; f += W) {
      const int n = isqrt_fix(f), m = f - n * n - n;
      const double2 R = Rt[n * n + n - m], dR = dRdn(Rt, n, -m, a.k, nrm);
      const double vr = -dR.x + se.x * R.x - se.y * R.y;
      const double vi = -dR.y + se.x * R.y + se.y * R.x;
      acc[f].x += -a.k * w * vi;
      acc[f].y += a.k * w * vr;
    }
    __syncthreads();
  }
  for (int f = lane; f < ps; f += W) {
    const int n = isqrt_fix(f);
    const double sg = (n & 1) ? -1.0 : 1.0;
    const double is = sg / a.wsc[n];
    const double2 B = make_double2(acc[f].x * is, acc[f].y * is);
#pragma unroll
    for (int r = 0; r < kR; r++)
      if (r < nr)
        col[r] = cadd2(col[r], cmul2(B, a.Mt[((size_t)box * a.nrhs + a.r0 + r) * ps + f]));
  }
  __syncthreads();
  double2 Qs = make_double2(0.0, 0.0);
#pragma unroll
  for (int r = 0; r < kR; r++) {
    if (r >= nr) continue;
    const double2 F = wsum2(fld[r]), Rr = wsum2(row[r]), Cr = wsum2(col[r]);
    if (lane == 0) {
      const int rr = a.r0 + r;
      const double2 xe = a.xs[(size_t)rr * N + e], le = a.lam[(size_t)rr * N + e];
      // b_e = e^{ik d.c} (1 + alpha i k d.n)
      const double *d = a.dirs + 3 * rr;
      const double dx = d[0] * c[0] + d[1] * c[1] + d[2] * c[2];
      const double dn = d[0] * nrm[0] + d[1] * nrm[1] + d[2] * nrm[2];
      double sn, cs;
      sincos(a.k * dx, &sn, &cs);
      const double fr = 1.0 - a.alim * a.k * dn, fi = a.alre * a.k * dn;
      const double2 b = make_double2(cs * fr - sn * fi, cs * fi + sn * fr);
      double2 t = cmul2(F, xe);
      t = cadd2(t, cmul2(le, b));
      const double2 lr = cmul2(le, Rr), cx = cmul2(Cr, xe);
      t = make_double2(t.x - lr.x - cx.x, t.y - lr.y - cx.y);
      t = cadd2(t, cmul2(cmul2(le, aee_s), xe));
      Qs = cadd2(Qs, t);
    }
  }
  if (lane == 0) {
    double2 *dst = a.Q + 18 * (size_t)e + s;
    *dst = cadd2(*dst, Qs);
  }
}

// dJ/dV[v][a] += 2 Re dQ_e/dV over the elements e of v
__global__ void k_shape_grad(const double2 *__restrict__ Q, const int *__restrict__ fv,
                             const double *__restrict__ geo, int N, double hrel,
                             double *__restrict__ dV) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N) return;
  const double h = hrel * geo[8 * (size_t)e + 7];
  for (int pi = 0; pi < 9; pi++) {
    const double2 qp = Q[18 * (size_t)e + 2 * pi], qm = Q[18 * (size_t)e + 2 * pi + 1];
    const double d = 2.0 * (qp.x - qm.x) / (2.0 * h);
    atomicAdd(dV + 3 * (size_t)fv[3 * (size_t)e + pi / 3] + pi % 3, d);
  }
}

__global__ void k_normal_part(const double *__restrict__ dV, const double *__restrict__ vn,
                              int nv, double *__restrict__ dvn) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  dvn[v] = dV[3 * v] * vn[3 * v] + dV[3 * v + 1] * vn[3 * v + 1] + dV[3 * v + 2] * vn[3 * v + 2];
}

__global__ void k_to_sorted(const int *__restrict__ perm, const double2 *__restrict__ x,
                            int N, int nrhs, double2 *__restrict__ xs) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)N * nrhs) return;
  const int r = (int)(i / N), e = (int)(i % N);
  xs[i] = x[(size_t)r * N + perm[e]];
}

inline int nblk(size_t n, int b = 256) { return (int)((n + b - 1) / b); }

}  // namespace

// x, lam: caller order [nrhs][N] (FP64); Lx, Mt: leaf expansions of the
// FP64 pass; u: conj(p) [nrhs][npts]; d_pts, d_dirs on the device.  dV
// ([nvert][3], device) is overwritten; dvn ([nvert], device) when not NULL.
extern "C" int fmbh_shape_grad(const fmb_plan *pl, const fmb_c64 *x, const fmb_c64 *lam,
                               int nrhs, const void *Lx, const void *Mt,
                               const fmb_c64 *u, const double *d_pts, int npts,
                               const double *d_dirs, double *dV, double *dvn,
                               void *stream) {
  const fmb_tree *t = pl->tree;
  hipStream_t st = (hipStream_t)stream;
  const int N = t->nelem, L = t->lmax, pL = pl->p[L], nm = pL;
  const double hrel = 1e-5;
  size_t lds = sizeof(double) * (2 * (nm + 1) + (nm + 1) * (nm + 2) / 2 + 4) +
               sizeof(double2) * ((nm + 1) + (nm + 1) * (nm + 1) + pL * pL);
  if (lds > 64 * 1024)
    return fmb_set_error(FMB_ERR_RANGE, "leaf order %d too large for shape gradients", pL);
  double2 *xs = NULL, *ls = NULL, *Q = NULL;
  int rc = fmbh_malloc((void **)&xs, sizeof(double2) * (size_t)N * nrhs);
  if (!rc) rc = fmbh_malloc((void **)&ls, sizeof(double2) * (size_t)N * nrhs);
  if (!rc) rc = fmbh_malloc((void **)&Q, sizeof(double2) * 18 * (size_t)N);
  if (rc) goto done;
  hipLaunchKernelGGL(k_to_sorted, dim3(nblk((size_t)N * nrhs)), dim3(256), 0, st, t->d_perm,
                     (const double2 *)x, N, nrhs, xs);
  hipLaunchKernelGGL(k_to_sorted, dim3(nblk((size_t)N * nrhs)), dim3(256), 0, st, t->d_perm,
                     (const double2 *)lam, N, nrhs, ls);
  if (hipMemsetAsync(Q, 0, sizeof(double2) * 18 * (size_t)N, st) != hipSuccess ||
      hipMemsetAsync(dV, 0, sizeof(double) * 3 * (size_t)t->nvert, st) != hipSuccess) {
    rc = fmb_set_error(FMB_ERR_HIP, "shape gradient: memset failed");
    goto done;
  }
  {
    ShapeArgs a;
    a.N = N;
    a.nrhs = nrhs;
    a.npts = npts;
    a.L = L;
    a.pL = pL;
    a.k = pl->k;
    a.alre = pl->alpha_re;
    a.alim = pl->alpha_im;
    a.D = t->D;
    a.p0 = t->pmin[0];
    a.p1 = t->pmin[1];
    a.p2 = t->pmin[2];
    a.ratio = pl->prm.near_ratio;
    a.hrel = hrel;
    a.tri = t->d_tri;
    a.geo = t->d_geo;
    a.ebox = t->d_elem_box;
    a.keys = t->d_keys[L];
    a.bstart = t->d_box_start;
    a.nstart = t->d_near_start;
    a.nbox = t->d_near_box;
    a.robin = (const double2 *)pl->d_robin;
    a.quad = pl->d_quad;
    a.nq = pl->nq;
    a.quadn = pl->d_quadn;
    a.nqn = pl->nqn;
    a.edge = pl->d_edge;
    a.nedge = pl->prm.nedge;
    a.wsc = pl->d_scale[L];
    a.xs = xs;
    a.lam = ls;
    a.Lx = (const double2 *)Lx;
    a.Mt = (const double2 *)Mt;
    a.pts = d_pts;
    a.u = (const double2 *)u;
    a.dirs = d_dirs;
    a.Q = Q;
    for (int r0 = 0; r0 < nrhs; r0 += kR) {
      a.r0 = r0;
      hipLaunchKernelGGL(k_shape_q, dim3(N, 18), dim3(W), lds, st, a);
      if (hipGetLastError() != hipSuccess) {
        rc = fmb_set_error(FMB_ERR_HIP, "shape gradient: launch failed");
        goto done;
      }
    }
  }
  hipLaunchKernelGGL(k_shape_grad, dim3(nblk(N)), dim3(256), 0, st, Q, t->d_fv, t->d_geo, N,
                     hrel, dV);
  if (dvn)
    hipLaunchKernelGGL(k_normal_part, dim3(nblk(t->nvert)), dim3(256), 0, st, dV, t->d_vnorm,
                       t->nvert, dvn);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
    rc = fmb_set_error(FMB_ERR_HIP, "shape gradient: kernel failed");
done:
  fmbh_free(xs);
  fmbh_free(ls);
  fmbh_free(Q);
  return rc;
}
